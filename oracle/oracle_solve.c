/*
 * oracle_solve.c -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * SpMV, BLAS-1, Jacobi, CG, GMRES, Backward Euler, Newton and the adjoint
 * driver, restated serially from the reference.  ORC_ORDER_GPU swaps each
 * dot product for the fixed block-tree reduction the HIP kernels use:
 *   - terms t_i, blocks of 256 consecutive indices, zero padded;
 *   - in a block: four 64-lane trees (lane l += lane l+off, off = 32..1),
 *     then the four wave sums added left to right;
 *   - across B blocks: lane k (k < 1024) adds block sums k, k+1024, ...
 *     left to right from 0.0; the 1024 lane sums go through sixteen 64-lane
 *     trees and the sixteen wave sums are added left to right.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

void orc__gather(const orc_problem* P, int64_t e, int mode, double time, const double* U, double* Ue);
void orc__element_matrix(const orc_problem* P, int64_t e, double time, int32_t n_forms,
                         const int32_t* kinds, const double* coefs, const double* U, double* Ke);

/* ------------------------------------------------------------------ */
/* reductions                                                           */
/* ------------------------------------------------------------------ */
/* nw 64-lane trees over v[64 * nw], wave sums added left to right */
static double wave_trees(double* v, int nw) {
    double r = 0.0;
    for (int k = 0; k < nw; ++k) {
        double* s = v + 64 * k;
        for (int off = 32; off >= 1; off >>= 1)
            for (int l = 0; l < off; ++l) s[l] = s[l] + s[l + off];
        r = k == 0 ? s[0] : r + s[0];
    }
    return r;
}

typedef double (*term_fn)(int64_t i, const void* ctx);

static double gpu_reduce(int64_t n, term_fn f, const void* ctx) {
    const int64_t B = (n + 255) / 256;
    double lane[1024];
    double v[256];
    for (int k = 0; k < 1024; ++k) lane[k] = 0.0;
    for (int64_t b = 0; b < B; ++b) {
        for (int t = 0; t < 256; ++t) {
            const int64_t i = b * 256 + t;
            v[t] = (i < n) ? f(i, ctx) : 0.0;
        }
        lane[b % 1024] += wave_trees(v, 4);
    }
    return wave_trees(lane, 16);
}

typedef struct { const double* x; const double* y; } xy_t;
static double term_xy(int64_t i, const void* c) {
    const xy_t* p = (const xy_t*)c;
    return p->x[i] * p->y[i];
}

/* VectorOps.tpp:7-18 (ref order) or the GPU block tree */
double orc_dot(int64_t n, const double* x, const double* y, int32_t order) {
    if (order == ORC_ORDER_GPU) {
        xy_t c = {x, y};
        return gpu_reduce(n, term_xy, &c);
    }
    double r = 0;
    for (int64_t i = 0; i < n; ++i) r += x[i] * y[i];
    return r;
}

/* MatrixOps.tpp:5-23 */
void orc_spmv(int64_t n, const int64_t* rp, const int64_t* ci, const double* val, const double* x, double* y) {
    for (int64_t i = 0; i < n; ++i) {
        double s = 0;
        for (int64_t p = rp[i]; p < rp[i + 1]; ++p) s += val[p] * x[ci[p]];
        y[i] = s;
    }
}

/* CSROperator::diagonal (CSROperator.hpp:40-45) + Jacobi::update (Jacobi.hpp:23-36) */
void orc_inv_diag(int64_t n, const int64_t* rp, const int64_t* ci, const double* val, double* d) {
    for (int64_t i = 0; i < n; ++i) {
        double v = 0;
        for (int64_t p = rp[i]; p < rp[i + 1]; ++p) if (ci[p] == i) v = val[p];
        d[i] = (fabs(v) > 1e-300) ? 1.0 / v : 1.0;
    }
}

static void precond_apply(int32_t pc, int64_t n, const double* invd, const double* r, double* z) {
    if (pc == 1) for (int64_t i = 0; i < n; ++i) z[i] = r[i] * invd[i];  /* VectorOps multiply */
    else for (int64_t i = 0; i < n; ++i) z[i] = r[i];                      /* Identity.hpp:15-25 */
}

/* ------------------------------------------------------------------ */
/* CG: linalg/solver/iterative/cg/Solver.tpp:15-119                     */
/* ------------------------------------------------------------------ */
/* The solver is written against an operator (apply) and a precomputed
 * Jacobi inverse diagonal, as cg::Solver is against its LinearOperator and
 * preconditioner concepts (linalg/operator/Operator.hpp:10-18). */
typedef void (*cg_apply_fn)(void* ctx, const double* x, double* y);

static int cg_core(int64_t n, cg_apply_fn apply, void* actx, const double* invd, const double* b, double* x,
                   double tol, int64_t max_iters, int32_t pc, int32_t order, orc_report* rep, double* hist) {
    double* r = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    double* p = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    double* Ap = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    double* z = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    apply(actx, x, Ap);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i];
    for (int64_t i = 0; i < n; ++i) r[i] += -1.0 * Ap[i];
    const double res0 = sqrt(orc_dot(n, r, r, order));
    rep->initial_residual = res0;
    int conv = 0;
    if (res0 / (res0 + 1e-50) < tol) {
        rep->converged = 1; rep->final_residual = res0; rep->final_residual_rel = 1.0;
        rep->iterations = 0; conv = 1;
        goto done;
    }
    precond_apply(pc, n, invd, r, z);
    for (int64_t i = 0; i < n; ++i) p[i] = z[i];
    double rz_old = orc_dot(n, r, z, order);
    for (int64_t k = 1; k < max_iters; ++k) {
        apply(actx, p, Ap);
        const double pAp = orc_dot(n, p, Ap, order);
        const double alpha = rz_old / pAp;
        for (int64_t i = 0; i < n; ++i) x[i] += alpha * p[i];
        for (int64_t i = 0; i < n; ++i) r[i] += -alpha * Ap[i];
        const double res = sqrt(orc_dot(n, r, r, order));
        const double rel = res / (res0 + 1e-50);
        if (hist) hist[k] = res;
        if (rel < tol) {
            rep->converged = 1; rep->final_residual = res; rep->final_residual_rel = rel;
            rep->iterations = k; conv = 1;
            goto done;
        }
        precond_apply(pc, n, invd, r, z);
        const double rz_new = orc_dot(n, r, z, order);
        const double beta = rz_new / rz_old;
        for (int64_t i = 0; i < n; ++i) p[i] *= beta;
        for (int64_t i = 0; i < n; ++i) p[i] += 1.0 * z[i];
        rz_old = rz_new;
    }
    {
        const double rf = sqrt(orc_dot(n, r, r, order));
        rep->converged = 0; rep->final_residual = rf; rep->final_residual_rel = rf / (res0 + 1e-50);
        rep->iterations = max_iters;
    }
done:
    if (hist) hist[0] = res0;
    free(r); free(p); free(Ap); free(z);
    return conv;
}

typedef struct { int64_t n; const int64_t* rp; const int64_t* ci; const double* val; } csr_op_t;

/* CSROperator::apply (CSROperator.hpp:26-29) */
static void csr_apply(void* c, const double* x, double* y) {
    const csr_op_t* A = (const csr_op_t*)c;
    orc_spmv(A->n, A->rp, A->ci, A->val, x, y);
}

int orc_cg(int64_t n, const int64_t* rp, const int64_t* ci, const double* val, const double* b,
           double* x, double tol, int64_t max_iters, int32_t pc, int32_t order, orc_report* rep,
           double* hist) {
    double* invd = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    if (pc == 1) orc_inv_diag(n, rp, ci, val, invd);
    csr_op_t A = {n, rp, ci, val};
    const int conv = cg_core(n, csr_apply, &A, invd, b, x, tol, max_iters, pc, order, rep, hist);
    free(invd);
    return conv;
}

typedef struct {
    const orc_problem* P; double time; int32_t nf; const int32_t* kinds; const double* coefs;
    const double* U; const double* Udot;
} fem_op_t;

/* FEMOperator::apply (FEMOperator.hpp:38-43): assembleVector with the
 * operand selector, the Krylov vector gathered from free DOFs only */
static void fem_apply(void* c, const double* x, double* y) {
    const fem_op_t* F = (const fem_op_t*)c;
    orc_assemble_vector(F->P, F->time, F->nf, F->kinds, F->coefs, F->U, F->Udot, x, ORC_SEL_OPERAND, y);
}

int orc_cg_fem(const orc_problem* P, double time, int32_t n_forms, const int32_t* kinds, const double* coefs,
               const double* U, const double* Udot, const double* b, double* x, double tol, int64_t max_iters,
               int32_t pc, int32_t order, orc_report* rep, double* hist) {
    const int64_t n = P->n_free;
    double* invd = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    if (pc == 1) {
        /* FEMOperator::diagonal (FEMOperator.hpp:49-51) + Jacobi::update (Jacobi.hpp:33-36) */
        orc_assemble_diagonal(P, time, n_forms, kinds, coefs, U, invd);
        for (int64_t i = 0; i < n; ++i) invd[i] = (fabs(invd[i]) > 1e-300) ? 1.0 / invd[i] : 1.0;
    }
    fem_op_t F = {P, time, n_forms, kinds, coefs, U, Udot};
    const int conv = cg_core(n, fem_apply, &F, invd, b, x, tol, max_iters, pc, order, rep, hist);
    free(invd);
    return conv;
}

/* ------------------------------------------------------------------ */
/* GMRES(m): linalg/solver/iterative/gmres/Solver.tpp:5-170             */
/* ------------------------------------------------------------------ */
int orc_gmres(int64_t n, const int64_t* rp, const int64_t* ci, const double* val, const double* b,
              double* x, double tol, int64_t max_iters, int64_t m, int32_t pc, int32_t order,
              orc_report* rep) {
    double** Q = (double**)malloc(sizeof(double*) * (size_t)(m + 1));
    for (int64_t i = 0; i <= m; ++i) Q[i] = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    double* w = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    double* z = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    double* invd = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    double* H = (double*)calloc((size_t)((m + 1) * m), sizeof(double));
    double* cs = (double*)calloc((size_t)(m + 1), sizeof(double));
    double* sn = (double*)calloc((size_t)(m + 1), sizeof(double));
    double* g = (double*)calloc((size_t)(m + 1), sizeof(double));
    double* y = (double*)calloc((size_t)(m + 1), sizeof(double));
#define HAT(i, j) H[(size_t)(j) * (size_t)(m + 1) + (size_t)(i)]
    if (pc == 1) orc_inv_diag(n, rp, ci, val, invd);
    const double eps = 1e-50, bdt = 1e-14;
    orc_spmv(n, rp, ci, val, x, w);
    for (int64_t i = 0; i < n; ++i) Q[0][i] = b[i];
    for (int64_t i = 0; i < n; ++i) Q[0][i] += -1.0 * w[i];
    const double res0 = sqrt(orc_dot(n, Q[0], Q[0], order));
    rep->initial_residual = res0;
    int result = 0;
    if (res0 / (res0 + eps) < tol) {
        rep->converged = 1; rep->final_residual = res0; rep->final_residual_rel = 1.0;
        rep->iterations = 0; result = 1; goto out;
    }
    int64_t total = 0;
    double beta = res0;
    while (total < max_iters) {
        const double sb = 1.0 / beta;
        for (int64_t i = 0; i < n; ++i) Q[0][i] *= sb;
        g[0] = beta;
        for (int64_t i = 1; i <= m; ++i) g[i] = 0.0;
        int64_t mloc = 0;
        int breakdown = 0;
        for (int64_t j = 0; j < m; ++j) {
            precond_apply(pc, n, invd, Q[j], z);
            orc_spmv(n, rp, ci, val, z, w);
            for (int64_t i = 0; i <= j; ++i) {
                const double hij = orc_dot(n, Q[i], w, order);
                HAT(i, j) = hij;
                for (int64_t k = 0; k < n; ++k) w[k] += -hij * Q[i][k];
            }
            const double hnext = sqrt(orc_dot(n, w, w, order));
            ++total;
            breakdown = (hnext <= bdt);
            if (!breakdown) {
                for (int64_t k = 0; k < n; ++k) Q[j + 1][k] = w[k];
                const double s = 1.0 / hnext;
                for (int64_t k = 0; k < n; ++k) Q[j + 1][k] *= s;
            }
            for (int64_t i = 0; i < j; ++i) {
                const double h1 = HAT(i, j), h2 = HAT(i + 1, j);
                HAT(i, j) = cs[i] * h1 + sn[i] * h2;
                HAT(i + 1, j) = -sn[i] * h1 + cs[i] * h2;
            }
            const double hjj = HAT(j, j);
            const double den = sqrt(hjj * hjj + hnext * hnext);
            if (den <= bdt) { cs[j] = 1.0; sn[j] = 0.0; }
            else { cs[j] = hjj / den; sn[j] = hnext / den; }
            HAT(j, j) = cs[j] * hjj + sn[j] * hnext;
            const double gj = g[j];
            g[j] = cs[j] * gj;
            g[j + 1] = -sn[j] * gj;
            const double est = fabs(g[j + 1]);
            mloc = j + 1;
            if (est / (res0 + eps) < tol || breakdown || total >= max_iters) break;
        }
        for (int64_t ii = 0; ii < mloc; ++ii) {
            const int64_t i = mloc - 1 - ii;
            double s = g[i];
            for (int64_t k = i + 1; k < mloc; ++k) s -= HAT(i, k) * y[k];
            y[i] = s / HAT(i, i);
        }
        for (int64_t k = 0; k < n; ++k) z[k] = 0.0;
        for (int64_t i = 0; i < mloc; ++i)
            for (int64_t k = 0; k < n; ++k) z[k] += y[i] * Q[i][k];
        precond_apply(pc, n, invd, z, w);
        for (int64_t k = 0; k < n; ++k) x[k] += 1.0 * w[k];
        orc_spmv(n, rp, ci, val, x, w);
        for (int64_t k = 0; k < n; ++k) Q[0][k] = b[k];
        for (int64_t k = 0; k < n; ++k) Q[0][k] += -1.0 * w[k];
        beta = sqrt(orc_dot(n, Q[0], Q[0], order));
        const int conv = beta / (res0 + eps) < tol;
        if (conv || breakdown || total >= max_iters) {
            rep->converged = conv; rep->final_residual = beta;
            rep->final_residual_rel = beta / (res0 + eps); rep->iterations = total;
            result = conv; goto out;
        }
    }
    rep->converged = 0; rep->final_residual = beta; rep->final_residual_rel = beta / (res0 + eps);
    rep->iterations = total;
out:
#undef HAT
    for (int64_t i = 0; i <= m; ++i) free(Q[i]);
    free(Q); free(w); free(z); free(invd); free(H); free(cs); free(sn); free(g); free(y);
    return result;
}

/* ------------------------------------------------------------------ */
/* drivers                                                              */
/* ------------------------------------------------------------------ */
static const int32_t K_DIFF[1] = {ORC_FORM_DIFFUSION};
static const double C_ONE[1] = {1.0};
static const int32_t K_MASS[1] = {ORC_FORM_MASS};

/* Linear BE: BackwardEulerRunner::step (BackwardEuler.hpp:132-172) driving
 * BackwardEulerStage::{setDt,assemble,solve,advance} (BackwardEulerStage.hpp:
 * 221-313).  The per-step residualNorm() of the reference is logging only
 * (ConstantStepSizePolicy ignores it) and is not reproduced. */
int64_t orc_be_run(const orc_problem* P, double t0, double tf, double dt_policy, const void* src,
                   const int64_t* rp, const int64_t* ci, const int64_t* map, int64_t nnz,
                   double cg_tol, int64_t cg_max, int32_t pc, int32_t order, double* U,
                   double* U_hist, int64_t n_steps_max, int64_t* iters) {
    const int64_t n = P->n_free;
    double* Uprev = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    double* Udot = (double*)calloc((size_t)(n + 1), sizeof(double));
    double* F = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    double* ms = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    double* val = (double*)malloc(sizeof(double) * (size_t)(nnz + 1));
    memcpy(Uprev, U, sizeof(double) * (size_t)n);
    double time = t0;
    int64_t step = 0;
    while (!(time >= tf - 1e-9)) {
        if (step >= n_steps_max) break;
        const double dt = (dt_policy < tf - time) ? dt_policy : (tf - time);
        const double ct = time + dt;
        const int32_t kinds[2] = {ORC_FORM_DIFFUSION, ORC_FORM_MASS | ORC_SCALED};
        const double coefs[2] = {1.0, 1.0 / dt};
        /* assemble() */
        for (int64_t i = 0; i < n; ++i) Udot[i] = U[i];
        const double a = -1.0 / dt, bb = 1.0 / dt;
        for (int64_t i = 0; i < n; ++i) Udot[i] = bb * Udot[i] + a * Uprev[i];
        orc_assemble_matrix(P, ct, 2, kinds, coefs, U, Udot, map, nnz, val);
        orc_assemble_load(P, ct, src, NULL, F);
        orc_assemble_vector(P, ct - dt, 1, K_MASS, C_ONE, Uprev, NULL, NULL, ORC_SEL_VALUE, ms);
        const double idt = 1.0 / dt;
        for (int64_t i = 0; i < n; ++i) F[i] += idt * ms[i];
        orc_apply_essential(P, ct, 2, kinds, coefs, F);
        /* solve(): CG warm-started from U */
        orc_report rep;
        orc_cg(n, rp, ci, val, F, U, cg_tol, cg_max, pc, order, &rep, NULL);
        if (iters) iters[step] = rep.iterations;
        if (!rep.converged) { step = -1; break; }
        /* advance() */
        memcpy(Uprev, U, sizeof(double) * (size_t)n);
        time += dt;
        if (U_hist) memcpy(U_hist + step * n, U, sizeof(double) * (size_t)n);
        ++step;
    }
    free(Uprev); free(Udot); free(F); free(ms); free(val);
    return step;
}

/* SteadyStage (linear mode): assemble (SteadyStage.hpp:60-72) + solve (:74-84) */
int orc_steady_linear(const orc_problem* P, const void* src, const double* nodal, const int64_t* rp,
                      const int64_t* ci, const int64_t* map, int64_t nnz, int32_t solver,
                      double tol, int64_t max_iters, int64_t krylov, int32_t pc, int32_t order,
                      double* U, orc_report* rep) {
    const int64_t n = P->n_free;
    double* val = (double*)malloc(sizeof(double) * (size_t)(nnz + 1));
    double* F = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    orc_assemble_matrix(P, 0.0, 1, K_DIFF, C_ONE, U, NULL, map, nnz, val);
    orc_assemble_load(P, 0.0, src, nodal, F);
    orc_apply_essential(P, 0.0, 1, K_DIFF, C_ONE, F);
    int ok = solver == 0 ? orc_cg(n, rp, ci, val, F, U, tol, max_iters, pc, order, rep, NULL)
                         : orc_gmres(n, rp, ci, val, F, U, tol, max_iters, krylov, pc, order, rep);
    free(val); free(F);
    return ok;
}

/* NewtonRunner::solve (Newton.hpp:32-83) over SteadyStage in nonlinear mode:
 * assemble() = load only (SteadyStage.hpp:60-72), residualNorm() = ||F - K(U)U||
 * (:89-96), solveLinearStep() = reassemble J = K + K_T in place, dU = 0,
 * solve J dU = R, U += dU (:101-116). */
int64_t orc_newton_steady(const orc_problem* P, const void* src, const int64_t* rp,
                          const int64_t* ci, const int64_t* map, int64_t nnz, double abs_tol,
                          double rel_tol, int64_t max_newton, int32_t lin_solver, double lin_tol,
                          int64_t lin_max, int64_t krylov, int32_t pc, int32_t order, double* U,
                          double* res_hist, int64_t* lin_iters) {
    const int64_t n = P->n_free;
    double* F = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    double* R = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    double* dU = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    double* J = (double*)malloc(sizeof(double) * (size_t)(nnz + 1));
    const int32_t jk[2] = {ORC_FORM_DIFFUSION, ORC_FORM_TDIFFUSION};
    const double jc[2] = {1.0, 1.0};
    double res0 = 0.0;
    int64_t result = -1;
    for (int64_t it = 0; it < max_newton; ++it) {
        orc_assemble_load(P, 0.0, src, NULL, F);
        orc_assemble_vector(P, 0.0, 1, K_DIFF, C_ONE, U, NULL, NULL, ORC_SEL_VALUE, R);
        for (int64_t i = 0; i < n; ++i) R[i] = -1.0 * R[i] + 1.0 * F[i];
        const double rn = sqrt(orc_dot(n, R, R, order));
        if (res_hist) res_hist[it] = rn;
        if (it == 0) res0 = rn;
        const double rr = (res0 > 0.0) ? rn / res0 : 0.0;
        if (rn < abs_tol || rr < rel_tol) { result = it; break; }
        orc_assemble_matrix(P, 0.0, 2, jk, jc, U, NULL, map, nnz, J);
        for (int64_t i = 0; i < n; ++i) dU[i] = 0.0;
        orc_report rep;
        const int ok = lin_solver == 0 ? orc_cg(n, rp, ci, J, R, dU, lin_tol, lin_max, pc, order, &rep, NULL)
                                       : orc_gmres(n, rp, ci, J, R, dU, lin_tol, lin_max, krylov, pc, order, &rep);
        if (lin_iters) lin_iters[it] = rep.iterations;
        if (!ok) { result = -2; break; }
        for (int64_t i = 0; i < n; ++i) U[i] += 1.0 * dU[i];
    }
    free(F); free(R); free(dU); free(J);
    return result;
}

/* Discrete adjoint of a fixed-dt linear BE run (no reference counterpart).
 * Forward: A u^{k+1} = F^{k+1} with A = K(kappa) + M/dt, checkpoints u^k.
 * Backward: A lam_N = dJ/du_N, A lam_k = (1/dt) M lam_{k+1};
 * dJ/dtheta_e = -sum_k kappa_e * lam_e^T Khat_e T_e^k, Khat_e = K_e at kappa = 1,
 * accumulated for k = N..1 in that order. */
double orc_adjoint_grad(const orc_problem* P, double t0, double dt, int64_t n_steps,
                        const int64_t* rp, const int64_t* ci, const int64_t* map, int64_t nnz,
                        double cg_tol, int64_t cg_max, int32_t order, const double* U0,
                        int64_t n_obs, const int64_t* obs, const double* t_obs, double* grad,
                        double* U_final) {
    const int64_t n = P->n_free, npe = P->npe;
    double* Uh = (double*)malloc(sizeof(double) * (size_t)((n_steps + 1) * n + 1));
    double* F = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    double* ms = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    double* A = (double*)malloc(sizeof(double) * (size_t)(nnz + 1));
    double* M = (double*)malloc(sizeof(double) * (size_t)(nnz + 1));
    double* lam = (double*)calloc((size_t)(n + 1), sizeof(double));
    double* rhs = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    const int32_t kinds[2] = {ORC_FORM_DIFFUSION, ORC_FORM_MASS | ORC_SCALED};
    const double coefs[2] = {1.0, 1.0 / dt};
    const double idt = 1.0 / dt;
    memcpy(Uh, U0, sizeof(double) * (size_t)n);
    orc_assemble_matrix(P, t0, 2, kinds, coefs, U0, NULL, map, nnz, A);
    orc_assemble_matrix(P, t0, 1, K_MASS, C_ONE, U0, NULL, map, nnz, M);
    double time = t0;
    for (int64_t k = 0; k < n_steps; ++k) {
        const double ct = time + dt;
        double* Uk = Uh + k * n;
        double* Un = Uh + (k + 1) * n;
        orc_assemble_vector(P, ct - dt, 1, K_MASS, C_ONE, Uk, NULL, NULL, ORC_SEL_VALUE, ms);
        for (int64_t i = 0; i < n; ++i) F[i] = idt * ms[i];
        orc_apply_essential(P, ct, 2, kinds, coefs, F);
        memcpy(Un, Uk, sizeof(double) * (size_t)n);
        orc_report rep;
        orc_cg(n, rp, ci, A, F, Un, cg_tol, cg_max, 1, order, &rep, NULL);
        time = ct;
    }
    const double* UN = Uh + n_steps * n;
    if (U_final) memcpy(U_final, UN, sizeof(double) * (size_t)n);
    double loss = 0.0;
    for (int64_t i = 0; i < n; ++i) rhs[i] = 0.0;
    for (int64_t k = 0; k < n_obs; ++k) {
        const double d = UN[obs[k]] - t_obs[k];
        loss += d * d;
        rhs[obs[k]] += d;  /* a repeated index contributes once per observation, as in the loss */
    }
    loss = 0.5 * loss;
    /* the element diffusion matrix at unit per-element factor: kappa_e is
     * taken out, the model's k_scalar and tensor stay */
    orc_problem Ph = *P;
    Ph.kappa_e = NULL;
    double Ke[64 * 64], Te[64], Le[64];
    for (int64_t e = 0; e < P->n_elem; ++e) grad[e] = 0.0;
    for (int64_t k = n_steps; k >= 1; --k) {
        if (k < n_steps) {
            orc_spmv(n, rp, ci, M, lam, F);
            for (int64_t i = 0; i < n; ++i) rhs[i] = idt * F[i];
        }
        for (int64_t i = 0; i < n; ++i) lam[i] = 0.0;
        orc_report rep;
        orc_cg(n, rp, ci, A, rhs, lam, cg_tol, cg_max, 1, order, &rep, NULL);
        const double tk = t0 + (double)k * dt;
        for (int64_t e = 0; e < P->n_elem; ++e) {
            orc__element_matrix(&Ph, e, tk, 1, K_DIFF, C_ONE, NULL, Ke);
            orc__gather(P, e, 1, tk, Uh + k * n, Te);
            orc__gather(P, e, 0, tk, lam, Le);
            double s = 0.0;
            for (int64_t a = 0; a < npe; ++a) {
                double v = 0.0;
                for (int64_t b = 0; b < npe; ++b) v += Ke[a * npe + b] * Te[b];
                s += Le[a] * v;
            }
            grad[e] += -(P->kappa_e[e] * s);
        }
    }
    free(Uh); free(F); free(ms); free(A); free(M); free(lam); free(rhs);
    return loss;
}
