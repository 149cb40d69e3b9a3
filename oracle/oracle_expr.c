/*
 * oracle_expr.c -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * A small recursive-descent expression evaluator standing in for the
 * reference's exprtk call sites (src/utils/expression/ScalarExpression.cpp:
 * 3-30, equation/heateq/evaluator/TemperatureExpression.hpp:38-51).
 * Variables: t, x, y, z, T.  exprtk itself is an un-vendored dependency
 * (downloaded by the reference CI at run time, no version pin), so its
 * arithmetic is "parity unpinned" beyond the expressions the reference
 * tests use ("1.0 + 0.5*T", "0.5", polynomial boundary data).
 *
 * Also holds the seeded Philox4x32-10 generator used for synthetic
 * per-element conductivity fields.
 */
#include "oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { OP_CONST, OP_VAR, OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_POW, OP_NEG, OP_FN1, OP_FN2 };
enum { F_SIN, F_COS, F_TAN, F_EXP, F_LOG, F_SQRT, F_ABS, F_POW, F_MIN, F_MAX };

typedef struct { int op; int arg; double c; } ins_t;
typedef struct { ins_t* code; int n, cap; int err; const char* s; } prog_t;

static void emit(prog_t* p, int op, int arg, double c) {
    if (p->n == p->cap) {
        p->cap = p->cap ? 2 * p->cap : 32;
        p->code = (ins_t*)realloc(p->code, sizeof(ins_t) * (size_t)p->cap);
    }
    p->code[p->n].op = op; p->code[p->n].arg = arg; p->code[p->n].c = c; p->n++;
}

static void skip_ws(prog_t* p) { while (*p->s && isspace((unsigned char)*p->s)) p->s++; }

static void parse_expr(prog_t* p);
static void parse_unary(prog_t* p);

static void parse_primary(prog_t* p) {
    skip_ws(p);
    const char* s = p->s;
    if (isdigit((unsigned char)*s) || (*s == '.' && isdigit((unsigned char)s[1]))) {
        char* end; double v = strtod(s, &end);
        p->s = end; emit(p, OP_CONST, 0, v); return;
    }
    if (isalpha((unsigned char)*s) || *s == '_') {
        char name[32]; int k = 0;
        while ((isalnum((unsigned char)*s) || *s == '_') && k < 31) name[k++] = *s++;
        name[k] = 0; p->s = s; skip_ws(p);
        if (*p->s == '(') {
            p->s++;
            static const char* f1[] = {"sin", "cos", "tan", "exp", "log", "sqrt", "abs"};
            static const char* f2[] = {"pow", "min", "max"};
            for (int i = 0; i < 7; ++i) if (!strcmp(name, f1[i])) {
                parse_expr(p); skip_ws(p);
                if (*p->s != ')') { p->err = 1; return; }
                p->s++; emit(p, OP_FN1, i, 0); return;
            }
            for (int i = 0; i < 3; ++i) if (!strcmp(name, f2[i])) {
                parse_expr(p); skip_ws(p);
                if (*p->s != ',') { p->err = 1; return; }
                p->s++; parse_expr(p); skip_ws(p);
                if (*p->s != ')') { p->err = 1; return; }
                p->s++; emit(p, OP_FN2, F_POW + i, 0); return;
            }
            p->err = 1; return;
        }
        if (!strcmp(name, "t")) { emit(p, OP_VAR, 0, 0); return; }
        if (!strcmp(name, "x")) { emit(p, OP_VAR, 1, 0); return; }
        if (!strcmp(name, "y")) { emit(p, OP_VAR, 2, 0); return; }
        if (!strcmp(name, "z")) { emit(p, OP_VAR, 3, 0); return; }
        if (!strcmp(name, "T")) { emit(p, OP_VAR, 4, 0); return; }
        if (!strcmp(name, "pi")) { emit(p, OP_CONST, 0, 3.14159265358979323846); return; }
        if (!strcmp(name, "e")) { emit(p, OP_CONST, 0, 2.71828182845904523536); return; }
        p->err = 1; return;
    }
    if (*s == '(') {
        p->s++; parse_expr(p); skip_ws(p);
        if (*p->s != ')') { p->err = 1; return; }
        p->s++; return;
    }
    p->err = 1;
}

/* power is right-associative and binds tighter than unary minus */
static void parse_power(prog_t* p) {
    parse_primary(p); skip_ws(p);
    if (*p->s == '^') { p->s++; parse_unary(p); emit(p, OP_POW, 0, 0); }
}

static void parse_unary(prog_t* p) {
    skip_ws(p);
    if (*p->s == '-') { p->s++; parse_unary(p); emit(p, OP_NEG, 0, 0); return; }
    if (*p->s == '+') { p->s++; parse_unary(p); return; }
    parse_power(p);
}

static void parse_term(prog_t* p) {
    parse_unary(p);
    for (;;) {
        skip_ws(p);
        if (*p->s == '*') { p->s++; parse_unary(p); emit(p, OP_MUL, 0, 0); }
        else if (*p->s == '/') { p->s++; parse_unary(p); emit(p, OP_DIV, 0, 0); }
        else return;
    }
}

static void parse_expr(prog_t* p) {
    parse_term(p);
    for (;;) {
        skip_ws(p);
        if (*p->s == '+') { p->s++; parse_term(p); emit(p, OP_ADD, 0, 0); }
        else if (*p->s == '-') { p->s++; parse_term(p); emit(p, OP_SUB, 0, 0); }
        else return;
    }
}

void* orc_expr_compile(const char* text) {
    prog_t* p = (prog_t*)calloc(1, sizeof(prog_t));
    p->s = text;
    parse_expr(p);
    skip_ws(p);
    if (p->err || *p->s) { free(p->code); free(p); return NULL; }
    return p;
}

void orc_expr_free(void* e) {
    prog_t* p = (prog_t*)e;
    if (!p) return;
    free(p->code); free(p);
}

double orc_expr_eval(const void* e, double t, double x, double y, double z, double T) {
    const prog_t* p = (const prog_t*)e;
    double st[64]; int sp = 0;
    const double var[5] = {t, x, y, z, T};
    for (int i = 0; i < p->n; ++i) {
        const ins_t* in = &p->code[i];
        switch (in->op) {
        case OP_CONST: st[sp++] = in->c; break;
        case OP_VAR: st[sp++] = var[in->arg]; break;
        case OP_ADD: sp--; st[sp - 1] = st[sp - 1] + st[sp]; break;
        case OP_SUB: sp--; st[sp - 1] = st[sp - 1] - st[sp]; break;
        case OP_MUL: sp--; st[sp - 1] = st[sp - 1] * st[sp]; break;
        case OP_DIV: sp--; st[sp - 1] = st[sp - 1] / st[sp]; break;
        case OP_POW: sp--; st[sp - 1] = pow(st[sp - 1], st[sp]); break;
        case OP_NEG: st[sp - 1] = -st[sp - 1]; break;
        case OP_FN1: {
            double a = st[sp - 1];
            switch (in->arg) {
            case F_SIN: a = sin(a); break;
            case F_COS: a = cos(a); break;
            case F_TAN: a = tan(a); break;
            case F_EXP: a = exp(a); break;
            case F_LOG: a = log(a); break;
            case F_SQRT: a = sqrt(a); break;
            default: a = fabs(a); break;
            }
            st[sp - 1] = a; break;
        }
        case OP_FN2: {
            sp--;
            double a = st[sp - 1], b = st[sp];
            if (in->arg == F_POW) a = pow(a, b);
            else if (in->arg == F_MIN) a = (b < a) ? b : a;
            else a = (b > a) ? b : a;
            st[sp - 1] = a; break;
        }
        }
    }
    return st[0];
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers:  */
/* as easy as 1, 2, 3", SC'11.  Counter = (i lo, i hi, stream, 0),       */
/* key = (seed lo, seed hi).                                            */
/* ------------------------------------------------------------------ */
static void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    philox(c, key[0], key[1]);
    for (int i = 0; i < 4; ++i) out[i] = c[i];
}

static double u53(uint32_t a, uint32_t b) {
    uint64_t v = (((uint64_t)a << 32) | b) >> 11;
    return ((double)v + 0.5) * (1.0 / 9007199254740992.0);
}

static void draw(uint64_t seed, uint32_t stream, int64_t i, double* u1, double* u2) {
    uint32_t c[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), stream, 0u};
    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    *u1 = u53(c[0], c[1]);
    *u2 = u53(c[2], c[3]);
}

void orc_normal_field(uint64_t seed, uint32_t stream, int64_t n, double* out) {
    for (int64_t i = 0; i < n; ++i) {
        double u1, u2;
        draw(seed, stream, i, &u1, &u2);
        out[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
}

void orc_uniform_field(uint64_t seed, uint32_t stream, int64_t n, double* out) {
    for (int64_t i = 0; i < n; ++i) {
        double u1, u2;
        draw(seed, stream, i, &u1, &u2);
        out[i] = u1;
    }
}
