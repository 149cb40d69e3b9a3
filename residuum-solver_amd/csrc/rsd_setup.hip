// rsd_setup.hip -- context, memory primitives, expressions and problem setup.
//
// Setup follows the reference's host-side steps: BlockMesh2D/3D::generate,
// TopologicalDOF::buildConstraints, Assembler<CPU>::createMatrix and the
// element->nnz positions scatterElementMatrix finds by merging.  The
// products are bitwise identical to the reference's (checked against the
// oracle and golden vectors); the pattern build is node-centric and
// multi-threaded rather than the reference's two-pass touch list.
#include "rsd_problem.hpp"
#include "rsd_assembly.hpp"
#include "rsd_linalg.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <thread>
#include <atomic>
#include <unordered_map>
#include <vector>

namespace rsd {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

// ------------------------------------------------------------------------
// expression compiler: recursive descent to a postfix program
// grammar: sum := prod (('+'|'-') prod)*, prod := unary (('*'|'/') unary)*,
// unary := '-' unary | '+' unary | pow, pow := atom ('^' unary)?
// ------------------------------------------------------------------------
namespace {
struct Parser {
    const char* s;
    Expr e;
    void emit(int op, double c = 0.0) {
        if (e.n >= kExprMax) throw Error(RSD_ERR_ARG, "expression too long");
        e.op[e.n] = op;
        e.c[e.n] = c;
        e.n++;
    }
    void ws() { while (*s && std::isspace((unsigned char)*s)) ++s; }
    [[noreturn]] void fail() { throw Error(RSD_ERR_ARG, std::string("cannot parse expression near '") + s + "'"); }
    void atom() {
        ws();
        if (std::isdigit((unsigned char)*s) || (*s == '.' && std::isdigit((unsigned char)s[1]))) {
            char* end;
            const double v = std::strtod(s, &end);
            s = end;
            emit(E_CONST, v);
            return;
        }
        if (std::isalpha((unsigned char)*s) || *s == '_') {
            std::string id;
            while (std::isalnum((unsigned char)*s) || *s == '_') id += *s++;
            ws();
            if (*s == '(') {
                ++s;
                struct F { const char* name; int op; int arity; };
                static const F fs[] = {{"sin", E_SIN, 1}, {"cos", E_COS, 1}, {"tan", E_TAN, 1},
                                       {"exp", E_EXP, 1}, {"log", E_LOG, 1}, {"sqrt", E_SQRT, 1},
                                       {"abs", E_ABS, 1}, {"pow", E_POW, 2}, {"min", E_MIN, 2},
                                       {"max", E_MAX, 2}};
                for (const F& f : fs) {
                    if (id != f.name) continue;
                    sum();
                    if (f.arity == 2) {
                        ws();
                        if (*s != ',') fail();
                        ++s;
                        sum();
                    }
                    ws();
                    if (*s != ')') fail();
                    ++s;
                    emit(f.op);
                    return;
                }
                fail();
            }
            static const char* vars[] = {"t", "x", "y", "z", "T"};
            for (int v = 0; v < 5; ++v)
                if (id == vars[v]) {
                    if (v == 0) e.uses_t = 1;
                    if (v == 4) e.uses_T = 1;
                    emit(E_VAR, (double)v);
                    return;
                }
            if (id == "pi") { emit(E_CONST, 3.14159265358979323846); return; }
            if (id == "e") { emit(E_CONST, 2.71828182845904523536); return; }
            fail();
        }
        if (*s == '(') {
            ++s;
            sum();
            ws();
            if (*s != ')') fail();
            ++s;
            return;
        }
        fail();
    }
    void power() {
        atom();
        ws();
        if (*s == '^') { ++s; unary(); emit(E_POW); }
    }
    void unary() {
        ws();
        if (*s == '-') { ++s; unary(); emit(E_NEG); return; }
        if (*s == '+') { ++s; unary(); return; }
        power();
    }
    void prod() {
        unary();
        for (;;) {
            ws();
            if (*s == '*') { ++s; unary(); emit(E_MUL); }
            else if (*s == '/') { ++s; unary(); emit(E_DIV); }
            else return;
        }
    }
    void sum() {
        prod();
        for (;;) {
            ws();
            if (*s == '+') { ++s; prod(); emit(E_ADD); }
            else if (*s == '-') { ++s; prod(); emit(E_SUB); }
            else return;
        }
    }
};
}  // namespace

Expr compile_expr(const std::string& text) {
    Parser p{text.c_str(), Expr{}};
    p.sum();
    p.ws();
    if (*p.s) p.fail();
    // stack depth check (device interpreter holds 16)
    int d = 0, dmax = 0;
    for (int i = 0; i < p.e.n; ++i) {
        const int op = p.e.op[i];
        if (op == E_CONST || op == E_VAR) ++d;
        else if (op == E_ADD || op == E_SUB || op == E_MUL || op == E_DIV || op == E_POW || op == E_MIN || op == E_MAX) --d;
        dmax = std::max(dmax, d);
    }
    if (dmax > 16) throw Error(RSD_ERR_ARG, "expression nests too deeply");
    return p.e;
}

double eval_expr_host(const Expr& e, double t, double x, double y, double z, double T) {
    double st[16];
    int sp = 0;
    for (int i = 0; i < e.n; ++i) {
        switch (e.op[i]) {
        case E_CONST: st[sp++] = e.c[i]; break;
        case E_VAR: { const int v = (int)e.c[i]; st[sp++] = v == 0 ? t : v == 1 ? x : v == 2 ? y : v == 3 ? z : T; break; }
        case E_ADD: sp--; st[sp - 1] = st[sp - 1] + st[sp]; break;
        case E_SUB: sp--; st[sp - 1] = st[sp - 1] - st[sp]; break;
        case E_MUL: sp--; st[sp - 1] = st[sp - 1] * st[sp]; break;
        case E_DIV: sp--; st[sp - 1] = st[sp - 1] / st[sp]; break;
        case E_POW: sp--; st[sp - 1] = std::pow(st[sp - 1], st[sp]); break;
        case E_MIN: sp--; st[sp - 1] = (st[sp] < st[sp - 1]) ? st[sp] : st[sp - 1]; break;
        case E_MAX: sp--; st[sp - 1] = (st[sp] > st[sp - 1]) ? st[sp] : st[sp - 1]; break;
        case E_NEG: st[sp - 1] = -st[sp - 1]; break;
        case E_SIN: st[sp - 1] = std::sin(st[sp - 1]); break;
        case E_COS: st[sp - 1] = std::cos(st[sp - 1]); break;
        case E_TAN: st[sp - 1] = std::tan(st[sp - 1]); break;
        case E_EXP: st[sp - 1] = std::exp(st[sp - 1]); break;
        case E_LOG: st[sp - 1] = std::log(st[sp - 1]); break;
        case E_SQRT: st[sp - 1] = std::sqrt(st[sp - 1]); break;
        default: st[sp - 1] = std::fabs(st[sp - 1]); break;
        }
    }
    return st[0];
}

// run f(lo, hi) over [0, n) on the host's cores
template<typename F>
static void parallel_for(int64_t n, F&& f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096) nt = 1;
    std::vector<std::thread> th;
    const int64_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const int64_t lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        th.emplace_back([&, lo, hi] { f(lo, hi); });
    }
    for (auto& x : th) x.join();
}

// ------------------------------------------------------------------------
// Dirichlet values per node
// ------------------------------------------------------------------------
__global__ void k_gval(int64_t n_nodes, int32_t dim, const double* __restrict__ xyz, const int32_t* __restrict__ t2a,
                       const int32_t* __restrict__ ctag, int32_t n_bc, const int32_t* __restrict__ tags,
                       const Expr* __restrict__ exprs, double time, double* __restrict__ g) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    double v = 0.0;
    if (t2a[i] < 0) {
        const double x = xyz[i * dim], y = xyz[i * dim + 1], z = dim == 3 ? xyz[i * dim + 2] : 0.0;
        for (int k = 0; k < n_bc; ++k)
            if (tags[k] == ctag[i]) v = eval_expr(exprs[k], time, x, y, z, 0.0);
    }
    g[i] = v;
}

// g[node[i]] = v[i]: caller-given Dirichlet values at the constrained nodes
__global__ void k_scatter_gval(int64_t n, const int32_t* __restrict__ node, const double* __restrict__ v,
                               double* __restrict__ g) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) g[node[i]] = v[i];
}

// SELL <-> CSR order for a problem's matrix, one thread per row: CSR entry
// j of row r sits at slot slice_ptr[r / 64] + 64 j + r % 64
__global__ void k_sell_csr(int64_t n, const int64_t* __restrict__ slice_ptr, const int32_t* __restrict__ rp,
                           double* __restrict__ sell, double* __restrict__ csr, int32_t to_csr) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t base = slice_ptr[r / kSlice] + r % kSlice;
    const int32_t b = rp[r], len = rp[r + 1] - b;
    for (int32_t j = 0; j < len; ++j) {
        if (to_csr) csr[b + j] = sell[base + (int64_t)j * kSlice];
        else sell[base + (int64_t)j * kSlice] = csr[b + j];
    }
}

void launch_sell_csr(const rsd_matrix* m, double* csr, bool to_csr) {
    const rsd_problem* pb = m->pb;
    if (pb->n_free)
        hipLaunchKernelGGL(k_sell_csr, dim3((unsigned)nblocks(pb->n_free)), dim3(kBlock), 0, pb->ctx->stream, pb->n_free,
                           pb->sell.slice_ptr.p, pb->sell.row_ptr.p, m->val.p, csr, to_csr ? 1 : 0);
    RSD_HIP(hipGetLastError());
}

// f(t, x, y, z) at the node of every free DOF, into the free vector
__global__ void k_eval_free(int64_t n_nodes, int32_t dim, const double* __restrict__ xyz,
                            const int32_t* __restrict__ t2a, Expr e, double time, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    const int32_t r = t2a[i];
    if (r < 0) return;
    const double x = xyz[i * dim], y = xyz[i * dim + 1], z = dim == 3 ? xyz[i * dim + 2] : 0.0;
    out[r] = eval_expr(e, time, x, y, z, 0.0);
}

}  // namespace rsd

using namespace rsd;

void rsd::update_gval(rsd_problem* pb, double time) {
    if (pb->gval_external) {
        RSD_CHECK(pb->gval_valid && pb->gval_time == time,
                  "the Dirichlet values at this time were not given (rsd_problem_set_dirichlet_values)");
        return;
    }
    if (pb->gval_valid && (!pb->bc_time_dependent || pb->gval_time == time)) return;
    const int64_t nb = nblocks(pb->n_nodes);
    hipLaunchKernelGGL(k_gval, dim3((unsigned)nb), dim3(kBlock), 0, pb->ctx->stream, pb->n_nodes, pb->dim,
                       pb->xyz.p, pb->t2a.p, pb->ctag.p, (int32_t)pb->bc_tags.size(), pb->d_bc_tags.p, pb->d_bc_exprs.p, time,
                       pb->gval.p);
    RSD_HIP(hipGetLastError());
    pb->gval_time = time;
    pb->gval_valid = true;
}

// ------------------------------------------------------------------------
// context
// ------------------------------------------------------------------------
double* rsd_ctx::partials(int64_t need) {
    if (need > scratch.partials_cap) {
        if (scratch.partials) RSD_HIP(hipFree(scratch.partials));
        const int64_t cap = std::max<int64_t>(need, 1 << 16);
        RSD_HIP(hipMalloc(&scratch.partials, sizeof(double) * (size_t)cap));
        scratch.partials_cap = cap;
    }
    return scratch.partials;
}

void* rsd_ctx::ws(const std::string& name, size_t bytes) {
    for (auto& e : ws_tab)
        if (e.first == name) {
            if (e.second.bytes < bytes) {
                RSD_HIP(hipStreamSynchronize(stream));
                RSD_HIP(hipFree(e.second.p));
                RSD_HIP(hipMalloc(&e.second.p, bytes));
                e.second.bytes = bytes;
            }
            return e.second.p;
        }
    Ws w;
    RSD_HIP(hipMalloc(&w.p, std::max<size_t>(bytes, 256)));
    w.bytes = std::max<size_t>(bytes, 256);
    ws_tab.emplace_back(name, w);
    return w.p;
}

rsd_ctx::~rsd_ctx() {
    for (auto& e : ws_tab) (void)hipFree(e.second.p);
}

// Sliced ELL layout of the CSR pattern in pb->h_rp / h_ci (rowlen[r] = row
// lengths), uploaded to the device; pb->n_free rows.
void rsd::build_sell(rsd_problem* pb, const std::vector<int32_t>& rowlen) {
    hipStream_t st = pb->ctx->stream;
    const int64_t n = pb->n_free;
    pb->max_row = n ? *std::max_element(rowlen.begin(), rowlen.end()) : 0;
    Sell& S = pb->sell;
    S.n_rows = n;
    S.n_slices = (n + kSlice - 1) / kSlice;
    S.h_slice_ptr.assign(S.n_slices + 1, 0);
    for (int64_t s = 0; s < S.n_slices; ++s) {
        int32_t w = 0;
        for (int64_t r = s * kSlice; r < std::min(n, (s + 1) * kSlice); ++r) w = std::max(w, rowlen[r]);
        S.h_slice_ptr[s + 1] = S.h_slice_ptr[s] + (int64_t)w * kSlice;
    }
    S.slots = S.h_slice_ptr[S.n_slices];
    RSD_CHECK(S.slots < (int64_t)INT32_MAX, "sliced layout exceeds int32 slots");
    std::vector<int32_t> hcol(S.slots);
    std::vector<int16_t> hoff(S.slots, 0);
    std::vector<int32_t> hdiag(n, -1);
    bool fits16 = true;
    for (int64_t s = 0; s < S.n_slices; ++s)
        for (int64_t r = s * kSlice; r < (s + 1) * kSlice; ++r) {
            const int64_t w = (S.h_slice_ptr[s + 1] - S.h_slice_ptr[s]) / kSlice;
            for (int64_t j = 0; j < w; ++j) {
                const int64_t sl = S.h_slice_ptr[s] + j * kSlice + (r % kSlice);
                int32_t c = r < n ? (int32_t)r : 0;
                if (r < n && j < rowlen[r]) c = pb->h_ci[pb->h_rp[r] + j];
                hcol[sl] = c;
                const int64_t off = (int64_t)c - (r < n ? r : 0);
                if (off < -32768 || off > 32767) fits16 = false;
                hoff[sl] = (int16_t)off;
                if (r < n && j < rowlen[r] && c == r) hdiag[r] = (int32_t)sl;
            }
        }
    S.has_off16 = fits16;
    // shared offset patterns: FNV-1a hash of (length, offsets), every hit
    // verified against the stored list; given up past 255 patterns, rows
    // longer than 64 or a table beyond 16 Ki entries
    {
        std::vector<std::vector<int32_t>> pats;
        std::unordered_map<uint64_t, int> by_hash;
        std::vector<uint8_t> pid(n);
        bool ok = pb->max_row <= 64;
        std::vector<int32_t> cur;
        for (int64_t r = 0; r < n && ok; ++r) {
            cur.clear();
            uint64_t h = 1469598103934665603ull ^ (uint64_t)rowlen[r];
            for (int32_t k = pb->h_rp[r]; k < pb->h_rp[r + 1]; ++k) {
                const int32_t o = pb->h_ci[k] - (int32_t)r;
                cur.push_back(o);
                h = (h ^ (uint64_t)(uint32_t)o) * 1099511628211ull;
            }
            auto it = by_hash.find(h);
            if (it != by_hash.end()) {
                if (pats[it->second] != cur) ok = false;
                else pid[r] = (uint8_t)it->second;
            } else if (pats.size() >= 255) {
                ok = false;
            } else {
                by_hash.emplace(h, (int)pats.size());
                pid[r] = (uint8_t)pats.size();
                pats.push_back(cur);
            }
        }
        const int32_t stride = std::max<int32_t>(pb->max_row, 1);
        ok = ok && n > 0 && (int64_t)pats.size() * stride <= 16384;
        S.has_pat = ok;
        if (ok) {
            std::vector<int32_t> off((size_t)pats.size() * stride, 0), len(pats.size());
            for (size_t q = 0; q < pats.size(); ++q) {
                len[q] = (int32_t)pats[q].size();
                std::copy(pats[q].begin(), pats[q].end(), off.begin() + (int64_t)q * stride);
            }
            S.n_pat = (int32_t)pats.size();
            S.pat_stride = stride;
            S.pat_id.upload(pid.data(), n, st);
            S.pat_off.upload(off.data(), (int64_t)off.size(), st);
            S.pat_len.upload(len.data(), (int64_t)len.size(), st);
            // stencil layout: the union of the patterns' offsets is 9 or 27
            // offsets in runs of three consecutive ones
            std::vector<int32_t> uni;
            for (const auto& q : pats) uni.insert(uni.end(), q.begin(), q.end());
            std::sort(uni.begin(), uni.end());
            uni.erase(std::unique(uni.begin(), uni.end()), uni.end());
            bool st_ok = uni.size() == 9 || uni.size() == 27;
            for (size_t k = 0; st_ok && k < uni.size(); k += 3)
                st_ok = uni[k + 1] == uni[k] + 1 && uni[k + 2] == uni[k] + 2;
            S.has_st = st_ok;
            if (st_ok) {
                S.st_np = (int32_t)uni.size();
                std::copy(uni.begin(), uni.end(), S.st_off);
                std::vector<uint32_t> pmask(pats.size(), 0u), rmask(n);
                for (size_t q = 0; q < pats.size(); ++q)
                    for (int32_t o : pats[q])
                        pmask[q] |= 1u << (std::lower_bound(uni.begin(), uni.end(), o) - uni.begin());
                for (int64_t r = 0; r < n; ++r) rmask[r] = pmask[pid[r]];
                S.st_mask.upload(rmask.data(), n, st);
                RSD_HIP(hipStreamSynchronize(st));  // rmask is a local
            }
            // slice-union layout: per slice the union of its rows' offsets
            const int64_t ns = S.n_slices;
            std::vector<int64_t> sptr(ns + 1, 0);
            std::vector<int32_t> soff((size_t)ns * kSuMax, 0), snu(ns, 0);
            std::vector<uint32_t> smask(n, 0u);
            bool su_ok = true;
            std::vector<int32_t> u;
            std::vector<uint8_t> seen(pats.size(), 0), ids;
            std::vector<uint32_t> pm(pats.size(), 0u);
            for (int64_t s = 0; s < ns && su_ok; ++s) {
                // the slice's distinct patterns (a few), then their offsets
                const int64_t r1 = std::min(n, (s + 1) * kSlice);
                ids.clear();
                for (int64_t r = s * kSlice; r < r1; ++r)
                    if (!seen[pid[r]]) {
                        seen[pid[r]] = 1;
                        ids.push_back(pid[r]);
                    }
                u.clear();
                for (uint8_t q : ids) {
                    seen[q] = 0;
                    u.insert(u.end(), pats[q].begin(), pats[q].end());
                }
                std::sort(u.begin(), u.end());
                u.erase(std::unique(u.begin(), u.end()), u.end());
                su_ok = u.size() <= (size_t)kSuMax;
                if (!su_ok) break;
                snu[s] = (int32_t)u.size();
                std::copy(u.begin(), u.end(), soff.begin() + s * kSuMax);
                sptr[s + 1] = sptr[s] + (int64_t)u.size() * kSlice;
                for (uint8_t q : ids) {
                    pm[q] = 0u;
                    for (int32_t o : pats[q]) pm[q] |= 1u << (std::lower_bound(u.begin(), u.end(), o) - u.begin());
                }
                for (int64_t r = s * kSlice; r < r1; ++r) smask[r] = pm[pid[r]];
            }
            su_ok = su_ok && sptr[ns] * 4 <= S.slots * 5 && sptr[ns] < (int64_t)INT32_MAX;
            S.has_su = su_ok;
            if (su_ok) {
                S.su_slots = sptr[ns];
                S.su_ptr.upload(sptr.data(), ns + 1, st);
                S.su_off.upload(soff.data(), (int64_t)soff.size(), st);
                S.su_nu.upload(snu.data(), ns, st);
                S.su_mask.upload(smask.data(), n, st);
                RSD_HIP(hipStreamSynchronize(st));  // the tables are locals
            }
        }
    }
    S.slice_ptr.upload(S.h_slice_ptr.data(), S.n_slices + 1, st);
    S.row_ptr.upload(pb->h_rp.data(), n + 1, st);
    S.col.upload(hcol.data(), S.slots, st);
    if (fits16) S.off16.upload(hoff.data(), S.slots, st);
    S.diag.upload(hdiag.data(), n, st);
}

// Symmetric half layout of the FP32 copy.  Row r's key is its list of
// (column - r, m) pairs in CSR order, m being the index of the pair's value
// in the upper list (columns >= row) of row min(r, c); rows with equal keys
// share one table entry.  Needs the shared offset patterns (rows of at most
// 64 entries) and a structurally symmetric pattern; given up past 255 keys.
void rsd::build_half(rsd_problem* pb) {
    Sell& S = pb->sell;
    if (S.half_tried) return;
    S.half_tried = true;
    if (!S.has_pat) return;
    const int64_t n = pb->n_free;
    const int32_t stride = S.pat_stride;
    const int32_t* rp = pb->h_rp.data();
    const int32_t* ci = pb->h_ci.data();
    // upper index of column r in row c (c < r), -1 when row c lacks it
    auto mirror = [&](int32_t c, int32_t r) -> int32_t {
        const int32_t* beg = ci + rp[c];
        const int32_t* end = ci + rp[c + 1];
        const int32_t* up = std::lower_bound(beg, end, c);
        const int32_t* it = std::lower_bound(up, end, r);
        return (it != end && *it == r) ? (int32_t)(it - up) : -1;
    };
    auto key_of = [&](int64_t r, std::vector<int32_t>& off, std::vector<int32_t>& mir) {
        off.clear();
        mir.clear();
        int32_t m_up = 0;
        for (int32_t k = rp[r]; k < rp[r + 1]; ++k) {
            const int32_t c = ci[k];
            off.push_back(c - (int32_t)r);
            mir.push_back(c < r ? mirror(c, (int32_t)r) : m_up++);
        }
    };
    auto hash_of = [](const std::vector<int32_t>& off, const std::vector<int32_t>& mir) {
        uint64_t h = 1469598103934665603ull ^ (uint64_t)off.size();
        for (size_t k = 0; k < off.size(); ++k) {
            h = (h ^ (uint64_t)(uint32_t)off[k]) * 1099511628211ull;
            h = (h ^ (uint64_t)(uint32_t)mir[k]) * 1099511628211ull;
        }
        return h;
    };
    // per-row hashes and upper counts in parallel; a missing mirror fails
    std::vector<uint64_t> hs(n);
    std::vector<int32_t> nup(n);
    std::atomic<bool> sym{true};
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        std::vector<int32_t> off, mir;
        for (int64_t r = lo; r < hi && sym.load(std::memory_order_relaxed); ++r) {
            key_of(r, off, mir);
            int32_t u = 0;
            for (size_t k = 0; k < off.size(); ++k) {
                if (mir[k] < 0) sym = false;
                u += off[k] >= 0;
            }
            nup[r] = u;
            hs[r] = hash_of(off, mir);
        }
    });
    if (!sym) return;
    // ids in first-occurrence order
    std::unordered_map<uint64_t, int> by_hash;
    std::vector<uint8_t> id(n);
    std::vector<int64_t> first;
    for (int64_t r = 0; r < n; ++r) {
        auto it = by_hash.find(hs[r]);
        if (it != by_hash.end()) { id[r] = (uint8_t)it->second; continue; }
        if (first.size() >= 255) return;
        by_hash.emplace(hs[r], (int)first.size());
        id[r] = (uint8_t)first.size();
        first.push_back(r);
    }
    const int64_t np = (int64_t)first.size();
    if (np * stride > 16384) return;
    std::vector<int32_t> toff(np * stride, 0), tmir(np * stride, 0), tlen(np), tlow(np);
    std::vector<int32_t> off, mir;
    for (int64_t q = 0; q < np; ++q) {
        key_of(first[q], off, mir);
        tlen[q] = (int32_t)off.size();
        tlow[q] = (int32_t)std::count_if(off.begin(), off.end(), [](int32_t o) { return o < 0; });
        std::copy(off.begin(), off.end(), toff.begin() + q * stride);
        std::copy(mir.begin(), mir.end(), tmir.begin() + q * stride);
    }
    // every row against its table entry (a hash collision gives up)
    std::atomic<bool> same{true};
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        std::vector<int32_t> o2, m2;
        for (int64_t r = lo; r < hi && same.load(std::memory_order_relaxed); ++r) {
            key_of(r, o2, m2);
            const int64_t q = id[r];
            if ((int32_t)o2.size() != tlen[q] || !std::equal(o2.begin(), o2.end(), toff.begin() + q * stride) ||
                !std::equal(m2.begin(), m2.end(), tmir.begin() + q * stride))
                same = false;
        }
    });
    if (!same) return;
    // ELL planes of n_pad values; every index below 2^31
    const int32_t width = n ? *std::max_element(nup.begin(), nup.end()) : 0;
    const int64_t pad = (n + kSlice - 1) / kSlice * kSlice;
    if (width == 0 || pad * (int64_t)width >= (int64_t)INT32_MAX) return;
    std::vector<int32_t> tval(np * stride, 0), tdst(np * stride, 0);
    for (int64_t q = 0; q < np; ++q)
        for (int32_t j = 0; j < tlen[q]; ++j) {
            const int32_t o = toff[q * stride + j];
            tval[q * stride + j] = (int32_t)(tmir[q * stride + j] * pad + std::min(o, 0));
            tdst[q * stride + j] = (int32_t)(tmir[q * stride + j] * pad);
        }
    // Uniform stencil: when the union of the rows' offsets is 3 or 9 runs of
    // 3 consecutive offsets, closed under negation, the values go to one
    // dense plane per non-negative offset (zeros where a row lacks it) and
    // the tables are rewritten to index those planes
    std::vector<int32_t> uni;
    for (int64_t q = 0; q < np; ++q) uni.insert(uni.end(), toff.begin() + q * stride, toff.begin() + q * stride + tlen[q]);
    std::sort(uni.begin(), uni.end());
    uni.erase(std::unique(uni.begin(), uni.end()), uni.end());
    bool ok = (uni.size() == 9 || uni.size() == 27);
    for (size_t k = 0; ok && k < uni.size(); k += 3) ok = uni[k + 1] == uni[k] + 1 && uni[k + 2] == uni[k] + 2;
    for (size_t k = 0; ok && k < uni.size(); ++k) ok = std::binary_search(uni.begin(), uni.end(), -uni[k]);
    std::unordered_map<int32_t, int32_t> plane;
    int64_t front = 0, ustride = 0;
    int32_t nplanes = 0;
    if (ok) {
        for (int32_t o : uni)
            if (o >= 0) plane[o] = nplanes++;
        front = -(int64_t)uni.front();
        ustride = (front + pad + kSlice - 1) / kSlice * kSlice;
        ok = ustride * nplanes < (int64_t)INT32_MAX;
    }
    auto vidx = [&](int32_t o) -> int32_t {  // value index - row of the entry with offset o
        return o >= 0 ? (int32_t)(plane[o] * ustride + front) : (int32_t)(plane[-o] * ustride + front + o);
    };
    if (ok) {
        for (int64_t q = 0; q < np; ++q)
            for (int32_t j = 0; j < tlen[q]; ++j) {
                const int32_t o = toff[q * stride + j];
                tval[q * stride + j] = vidx(o);
                tdst[q * stride + j] = o >= 0 ? vidx(o) : 0;
            }
        S.uni_nr = (int32_t)uni.size() / 3;
        for (int32_t k = 0; k < S.uni_nr; ++k) {
            S.uni_roff[k] = uni[3 * k];
            for (int i = 0; i < 3; ++i) S.uni_vidx[3 * k + i] = vidx(uni[3 * k + i]);
        }
        S.uni_planes = nplanes;
        S.uni_stride = ustride;
        S.uni_front = front;
        // a 12-byte load at r + roff reads up to r + max offset; pad to whole
        // 64-float groups so the vector's element 0 keeps its alignment
        S.uni_lo = (front + 4 + kSlice - 1) / kSlice * kSlice;
        S.uni_hi = ((int64_t)uni.back() + 4 + kSlice - 1) / kSlice * kSlice;
        S.has_uni = true;
    }
    hipStream_t st = pb->ctx->stream;
    S.half_dst.upload(tdst.data(), (int64_t)tdst.size(), st);
    S.half_id.upload(id.data(), n, st);
    S.half_off.upload(toff.data(), (int64_t)toff.size(), st);
    S.half_mir.upload(tmir.data(), (int64_t)tmir.size(), st);
    S.half_val.upload(tval.data(), (int64_t)tval.size(), st);
    S.half_len.upload(tlen.data(), np, st);
    S.half_low.upload(tlow.data(), np, st);
    S.n_half = (int32_t)np;
    S.half_width = width;
    S.half_pad = pad;
    S.has_half = true;
}

// BlockMesh2D/3D::generate (src/mesh/generator/BlockMesh2D.cpp:6-121,
// BlockMesh3D.cpp:6-139) with one basis order P in every direction
static void block_mesh(int32_t dim, const int64_t* nel, const double* ext, int64_t P, std::vector<double>& xyz,
                       std::vector<int32_t>& ien, std::vector<int32_t>& rng, int64_t* n_nodes, int64_t* n_elem) {
    const int64_t nx = nel[0], ny = nel[1], nz = dim == 3 ? nel[2] : 1;
    RSD_CHECK(nx > 0 && ny > 0 && nz > 0, "element counts must be positive");
    const int64_t NX = nx * P + 1, NY = ny * P + 1, NZ = dim == 3 ? nz * P + 1 : 1;
    const int64_t nn = NX * NY * NZ, ne = nx * ny * nz;
    RSD_CHECK(nn < INT32_MAX, "mesh too large for int32 node ids");
    *n_nodes = nn;
    *n_elem = ne;
    // nodes: x fastest (BlockMesh2D.cpp:50-71, BlockMesh3D.cpp:50-76)
    xyz.resize(nn * dim);
    const double dx = (ext[1] - ext[0]) / (double)(nx * P);
    const double dy = (ext[3] - ext[2]) / (double)(ny * P);
    const double dz = dim == 3 ? (ext[5] - ext[4]) / (double)(nz * P) : 0.0;
    int64_t node = 0;
    for (int64_t k = 0; k < NZ; ++k)
        for (int64_t j = 0; j < NY; ++j)
            for (int64_t i = 0; i < NX; ++i, ++node) {
                xyz[dim * node] = ext[0] + (double)i * dx;
                xyz[dim * node + 1] = ext[2] + (double)j * dy;
                if (dim == 3) xyz[dim * node + 2] = ext[4] + (double)k * dz;
            }
    // connectivity, local node a -> (a % (P+1), ...) (BlockMesh3D.cpp:78-108)
    const int npe = (int)(dim == 2 ? (P + 1) * (P + 1) : (P + 1) * (P + 1) * (P + 1));
    const int nfaces = 2 * dim;
    ien.resize(ne * npe);
    rng.resize(ne * nfaces);
    for (int64_t ez = 0; ez < nz; ++ez)
        for (int64_t ey = 0; ey < ny; ++ey)
            for (int64_t ex = 0; ex < nx; ++ex) {
                const int64_t e = (ez * ny + ey) * nx + ex;
                for (int a = 0; a < npe; ++a) {
                    const int64_t ax = a % (P + 1), ay = (a / (P + 1)) % (P + 1), az = a / ((P + 1) * (P + 1));
                    ien[e * npe + a] = (int32_t)(((ez * P + az) * NY + (ey * P + ay)) * NX + ex * P + ax);
                }
                int32_t* r = &rng[e * nfaces];
                r[0] = ex == 0 ? 0 : -1;
                r[1] = ex == nx - 1 ? 1 : -1;
                r[2] = ey == 0 ? 2 : -1;
                r[3] = ey == ny - 1 ? 3 : -1;
                if (dim == 3) {
                    r[4] = ez == 0 ? 4 : -1;
                    r[5] = ez == nz - 1 ? 5 : -1;
                }
            }
}

extern "C" {

const char* rsd_last_error_string(void) { return g_last_error.c_str(); }
const char* rsd_version(void) { return "residuum-hip 0.1 (gfx950)"; }

#ifndef RSD_ABI_HASH
#error "RSD_ABI_HASH (the SHA-256 prefix of include/residuum_hip.h) is set by the Makefile"
#endif
const char* rsd_abi_hash(void) { return RSD_ABI_HASH; }

int rsd_ctx_create(int device, rsd_ctx** out) {
    return guard([&] {
        RSD_CHECK(out, "rsd_ctx_create: null out");
        auto c = std::make_unique<rsd_ctx>();
        c->device = device;
        RSD_HIP(hipSetDevice(device));
        RSD_HIP(hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device));
        RSD_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        RSD_HIP(hipMalloc(&c->d_scalars, sizeof(double) * 64));
        c->partials(1 << 16);
        RSD_HIP(hipStreamSynchronize(c->stream));
        *out = c.release();
    });
}

int rsd_ctx_destroy(rsd_ctx* ctx) {
    return guard([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        if (ctx->scratch.partials) (void)hipFree(ctx->scratch.partials);
        if (ctx->d_scalars) (void)hipFree(ctx->d_scalars);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

int rsd_ctx_sync(rsd_ctx* ctx) {
    return guard([&] { RSD_CHECK(ctx, "null ctx"); RSD_HIP(hipStreamSynchronize(ctx->stream)); });
}

int rsd_ctx_set_gmres_orthogonalization(rsd_ctx* ctx, int32_t mode) {
    return guard([&] {
        RSD_CHECK(ctx, "null ctx");
        RSD_CHECK(mode == RSD_ORTHO_MGS || mode == RSD_ORTHO_CGS2, "unknown GMRES orthogonalisation");
        ctx->gmres_orth = mode;
    });
}

int rsd_ctx_stream(rsd_ctx* ctx, void** stream) {
    return guard([&] { RSD_CHECK(ctx && stream, "null argument"); *stream = (void*)ctx->stream; });
}

int rsd_malloc(rsd_ctx* ctx, size_t bytes, void** out) {
    return guard([&] {
        RSD_CHECK(ctx && out, "rsd_malloc: null argument");
        *out = nullptr;
        if (bytes) RSD_HIP(hipMalloc(out, bytes));
    });
}

int rsd_free(rsd_ctx* ctx, void* p) {
    return guard([&] {
        RSD_CHECK(ctx, "rsd_free: null ctx");
        if (p) {
            RSD_HIP(hipStreamSynchronize(ctx->stream));
            RSD_HIP(hipFree(p));
        }
    });
}

int rsd_memcpy(rsd_ctx* ctx, void* dst, const void* src, size_t bytes, int kind) {
    return guard([&] {
        RSD_CHECK(ctx, "rsd_memcpy: null ctx");
        RSD_CHECK(kind >= 0 && kind <= 3, "rsd_memcpy: bad copy kind");
        if (!bytes) return;
        RSD_CHECK(dst && src, "rsd_memcpy: null pointer");
        static const hipMemcpyKind k[4] = {hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost,
                                           hipMemcpyDeviceToDevice};
        RSD_HIP(hipMemcpyAsync(dst, src, bytes, k[kind], ctx->stream));
        if (kind == RSD_COPY_D2H || kind == RSD_COPY_H2H || kind == RSD_COPY_H2D) RSD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int rsd_memset_zero(rsd_ctx* ctx, void* p, size_t bytes) {
    return guard([&] {
        RSD_CHECK(ctx, "null ctx");
        if (bytes) {
            RSD_CHECK(p, "rsd_memset_zero: null pointer");
            RSD_HIP(hipMemsetAsync(p, 0, bytes, ctx->stream));
        }
    });
}

// ------------------------------------------------------------------------
// problem setup
// ------------------------------------------------------------------------
static void face_nodes(int dim, int p, int f, std::vector<int>& out) {
    out.clear();
    const int X = p, Y = p, Z = p;
    if (dim == 2) {
        if (f == 0) for (int i = 0; i <= Y; ++i) out.push_back(i * (X + 1));
        if (f == 1) for (int i = 0; i <= Y; ++i) out.push_back(i * (X + 1) + X);
        if (f == 2) for (int i = 0; i <= X; ++i) out.push_back(i);
        if (f == 3) for (int i = 0; i <= X; ++i) out.push_back(i + (X + 1) * Y);
        return;
    }
    const int XY = (X + 1) * (Y + 1);
    for (int j = 0; j <= (f < 4 ? Z : Y); ++j)
        for (int i = 0; i <= (f < 2 ? Y : X); ++i) {
            switch (f) {
            case 0: out.push_back(i * (X + 1) + j * XY); break;
            case 1: out.push_back(i * (X + 1) + j * XY + X); break;
            case 2: out.push_back(i + j * XY); break;
            case 3: out.push_back(i + j * XY + (X + 1) * Y); break;
            case 4: out.push_back(i + j * (X + 1)); break;
            default: out.push_back(i + j * (X + 1) + XY * Z); break;
            }
        }
}

// Row gather tables.  Free row r (slice r / 64, lane r % 64) lists the
// elements around its node in colour order -- distinct colours, since
// elements sharing a node never share one -- as e * npe + a (a: the node's
// local index), and for each local column b the position of column
// t2a(ien[e][b]) in the row's sorted column list (0xff when constrained).
// Summing the element rows in that order gives every entry its
// contributions in the colour order of the element scatter.  Not built
// when an element repeats a node, a row is wider than the gather kernel's
// accumulator or e * npe overflows int32.
static void build_gather(rsd_problem* pb, const std::vector<int64_t>& n2e_ptr, const std::vector<int32_t>& n2e) {
    pb->gat_m = 0;
    const int npe = pb->npe;
    const int64_t n = pb->n_free, ne = pb->n_elem;
    const Sell& S = pb->sell;
    if (n == 0 || ne == 0 || pb->max_row > kGatherMaxRow || ne * npe >= (int64_t)INT32_MAX) return;
    for (int64_t e = 0; e < ne; ++e)
        for (int a = 0; a < npe; ++a)
            for (int b = 0; b < a; ++b)
                if (pb->h_ien[e * npe + a] == pb->h_ien[e * npe + b]) return;
    int64_t M = 0;
    for (int64_t r = 0; r < n; ++r) {
        const int32_t v = pb->h_a2t[r];
        M = std::max(M, n2e_ptr[v + 1] - n2e_ptr[v]);
    }
    if (M == 0 || M > kGatherMaxElems) return;
    const int k4 = (npe + 3) / 4;
    std::vector<int32_t> gel((size_t)S.n_slices * M * kSlice, -1);
    std::vector<uint32_t> gk((size_t)S.n_slices * M * k4 * kSlice, 0xffffffffu);
    std::atomic<bool> ok{true};
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        std::vector<std::pair<int32_t, int32_t>> el;
        for (int64_t r = lo; r < hi; ++r) {
            const int32_t v = pb->h_a2t[r];
            el.clear();
            for (int64_t k = n2e_ptr[v]; k < n2e_ptr[v + 1]; ++k) el.emplace_back(pb->h_colour[n2e[k]], n2e[k]);
            std::sort(el.begin(), el.end());
            const int64_t s = r / kSlice, lane = r % kSlice;
            const int32_t* beg = pb->h_ci.data() + pb->h_rp[r];
            const int32_t* end = pb->h_ci.data() + pb->h_rp[r + 1];
            for (size_t j = 0; j < el.size(); ++j) {
                if (j && el[j].first == el[j - 1].first) ok = false;
                const int64_t e = el[j].second;
                int a = 0;
                while (a < npe && pb->h_ien[e * npe + a] != v) ++a;
                gel[((size_t)s * M + j) * kSlice + lane] = (int32_t)(e * npe + a);
                for (int b = 0; b < npe; ++b) {
                    const int32_t c = pb->h_t2a[pb->h_ien[e * npe + b]];
                    uint32_t kb = 0xffu;
                    if (c >= 0) kb = (uint32_t)(std::lower_bound(beg, end, c) - beg);
                    uint32_t& w = gk[(((size_t)s * M + j) * k4 + b / 4) * kSlice + lane];
                    w = (w & ~(0xffu << (8 * (b % 4)))) | (kb << (8 * (b % 4)));
                }
            }
        }
    });
    if (!ok) return;
    hipStream_t st = pb->ctx->stream;
    pb->gat_el.upload(gel.data(), (int64_t)gel.size(), st);
    pb->gat_k.upload(gk.data(), (int64_t)gk.size(), st);
    RSD_HIP(hipStreamSynchronize(st));
    pb->gat_m = (int32_t)M;
}

static void build_problem(rsd_problem* pb) {
    const int dim = pb->dim, npe = pb->npe;
    const int64_t ne = pb->n_elem, nn = pb->n_nodes;
    hipStream_t st = pb->ctx->stream;
    // constraints: TopologicalDOF::buildConstraints
    std::vector<char> con(nn, 0);
    pb->h_ctag.assign(nn, -1);
    std::vector<int> fn;
    for (int64_t e = 0; e < ne; ++e)
        for (int f = 0; f < pb->nfaces; ++f) {
            const int32_t tag = pb->h_rng[e * pb->nfaces + f];
            if (tag < 0) continue;
            if (std::find(pb->bc_tags.begin(), pb->bc_tags.end(), tag) == pb->bc_tags.end()) continue;
            face_nodes(dim, pb->p, f, fn);
            for (int a : fn) {
                const int32_t g = pb->h_ien[e * npe + a];
                con[g] = 1;
                pb->h_ctag[g] = tag;
            }
        }
    pb->h_t2a.assign(nn, -1);
    pb->h_a2t.clear();
    for (int64_t i = 0; i < nn; ++i)
        if (!con[i]) {
            pb->h_t2a[i] = (int32_t)pb->h_a2t.size();
            pb->h_a2t.push_back((int32_t)i);
        }
    const int64_t n = (int64_t)pb->h_a2t.size();
    pb->n_free = n;
    // node -> elements
    std::vector<int64_t> n2e_ptr(nn + 1, 0);
    for (int64_t k = 0; k < ne * npe; ++k) n2e_ptr[pb->h_ien[k] + 1]++;
    for (int64_t i = 0; i < nn; ++i) n2e_ptr[i + 1] += n2e_ptr[i];
    std::vector<int32_t> n2e(n2e_ptr[nn]);
    {
        std::vector<int64_t> cur(n2e_ptr.begin(), n2e_ptr.end() - 1);
        for (int64_t e = 0; e < ne; ++e)
            for (int a = 0; a < npe; ++a) n2e[cur[pb->h_ien[e * npe + a]]++] = (int32_t)e;
    }
    // pattern: row r = sorted unique free columns of elements containing node(r)
    std::vector<int32_t> rowlen(n);
    auto row_cols = [&](int64_t r, std::vector<int32_t>& cols) {
        cols.clear();
        const int32_t node = pb->h_a2t[r];
        for (int64_t k = n2e_ptr[node]; k < n2e_ptr[node + 1]; ++k) {
            const int64_t e = n2e[k];
            for (int b = 0; b < npe; ++b) {
                const int32_t c = pb->h_t2a[pb->h_ien[e * npe + b]];
                if (c >= 0) cols.push_back(c);
            }
        }
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    };
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        std::vector<int32_t> cols;
        for (int64_t r = lo; r < hi; ++r) { row_cols(r, cols); rowlen[r] = (int32_t)cols.size(); }
    });
    pb->h_rp.assign(n + 1, 0);
    for (int64_t r = 0; r < n; ++r) {
        RSD_CHECK((int64_t)pb->h_rp[r] + rowlen[r] < (int64_t)INT32_MAX, "pattern exceeds int32 nnz");
        pb->h_rp[r + 1] = pb->h_rp[r] + rowlen[r];
    }
    pb->nnz = pb->h_rp[n];
    pb->h_ci.resize(pb->nnz);
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        std::vector<int32_t> cols;
        for (int64_t r = lo; r < hi; ++r) {
            row_cols(r, cols);
            std::copy(cols.begin(), cols.end(), pb->h_ci.begin() + pb->h_rp[r]);
        }
    });
    build_sell(pb, rowlen);
    Sell& S = pb->sell;
    auto slot_of = [&](int64_t r, int64_t j) { return S.h_slice_ptr[r / kSlice] + j * kSlice + (r % kSlice); };
    // element -> SELL slot map
    std::vector<int32_t> hmap((size_t)ne * npe * npe);
    parallel_for(ne, [&](int64_t lo, int64_t hi) {
        for (int64_t e = lo; e < hi; ++e)
            for (int a = 0; a < npe; ++a) {
                const int32_t r = pb->h_t2a[pb->h_ien[e * npe + a]];
                for (int b = 0; b < npe; ++b) {
                    const int32_t c = pb->h_t2a[pb->h_ien[e * npe + b]];
                    int32_t v = -1;
                    if (r >= 0 && c >= 0) {
                        const int32_t* beg = pb->h_ci.data() + pb->h_rp[r];
                        const int32_t* end = pb->h_ci.data() + pb->h_rp[r + 1];
                        const int32_t* it = std::lower_bound(beg, end, c);
                        v = (int32_t)slot_of(r, it - beg);
                    }
                    hmap[((size_t)e * npe + a) * npe + b] = v;
                }
            }
    });
    pb->map_sell.upload(hmap.data(), (int64_t)hmap.size(), st);
    // first-fit greedy colouring in element order, colour-major order
    pb->h_colour.assign(ne, 0);
    std::vector<uint64_t> mask(nn, 0);
    int nc = 0;
    for (int64_t e = 0; e < ne; ++e) {
        uint64_t used = 0;
        for (int a = 0; a < npe; ++a) used |= mask[pb->h_ien[e * npe + a]];
        int c = 0;
        while (used & (1ull << c)) ++c;
        RSD_CHECK(c < 64, "colouring needs more than 64 colours");
        pb->h_colour[e] = c;
        nc = std::max(nc, c + 1);
        for (int a = 0; a < npe; ++a) mask[pb->h_ien[e * npe + a]] |= (1ull << c);
    }
    pb->colour_off.assign(nc + 1, 0);
    for (int64_t e = 0; e < ne; ++e) pb->colour_off[pb->h_colour[e] + 1]++;
    for (int c = 0; c < nc; ++c) pb->colour_off[c + 1] += pb->colour_off[c];
    pb->h_order.resize(ne);
    {
        std::vector<int64_t> cur(pb->colour_off.begin(), pb->colour_off.end() - 1);
        for (int64_t e = 0; e < ne; ++e) pb->h_order[cur[pb->h_colour[e]]++] = (int32_t)e;
    }
    // the elements touching a constrained node, in the same colour-major
    // order: the essential lift only adds exact zeros on the others
    {
        std::vector<int32_t> border;
        pb->bcolour_off.assign(nc + 1, 0);
        for (int c = 0; c < nc; ++c) {
            for (int64_t k = pb->colour_off[c]; k < pb->colour_off[c + 1]; ++k) {
                const int32_t e = pb->h_order[k];
                bool touches = false;
                for (int a = 0; a < npe; ++a) touches = touches || pb->h_t2a[pb->h_ien[(int64_t)e * npe + a]] < 0;
                if (touches) border.push_back(e);
            }
            pb->bcolour_off[c + 1] = (int64_t)border.size();
        }
        pb->border.upload(border.data(), (int64_t)border.size(), st);
        // the lift's row lists: entry k * npe + a of border element k goes
        // to free row t2a(ien[e][a]); the entries of a row in border order
        // are its contributions in colour order
        std::vector<std::pair<int32_t, int32_t>> ent;
        const bool fits = (int64_t)border.size() * npe < (int64_t)INT32_MAX;
        for (size_t k = 0; fits && k < border.size(); ++k)
            for (int a = 0; a < npe; ++a) {
                const int32_t r = pb->h_t2a[pb->h_ien[(int64_t)border[k] * npe + a]];
                if (r >= 0) ent.emplace_back(r, (int32_t)(k * npe + a));
            }
        std::stable_sort(ent.begin(), ent.end(),
                         [](const std::pair<int32_t, int32_t>& x, const std::pair<int32_t, int32_t>& y) {
                             return x.first < y.first;
                         });
        std::vector<int32_t> rows, off{0}, idx;
        for (size_t i = 0; i < ent.size(); ++i) {
            if (i == 0 || ent[i].first != ent[i - 1].first) {
                if (i) off.push_back((int32_t)idx.size());
                rows.push_back(ent[i].first);
            }
            idx.push_back(ent[i].second);
        }
        if (!rows.empty()) off.push_back((int32_t)idx.size());
        pb->lift_nrows = (int64_t)rows.size();
        pb->lift_row.upload(rows.data(), (int64_t)rows.size(), st);
        pb->lift_off.upload(off.data(), (int64_t)off.size(), st);
        pb->lift_idx.upload(idx.data(), (int64_t)idx.size(), st);
        pb->lift_buf.alloc((int64_t)border.size() * npe);
    }
    build_gather(pb, n2e_ptr, n2e);
    build_shape_table(pb);
    // device copies
    pb->xyz.upload(pb->h_xyz.data(), nn * dim, st);
    pb->ien.upload(pb->h_ien.data(), ne * npe, st);
    pb->t2a.upload(pb->h_t2a.data(), nn, st);
    pb->ctag.upload(pb->h_ctag.data(), nn, st);
    pb->order.upload(pb->h_order.data(), ne, st);
    pb->gval.alloc(nn);
    pb->d_bc_tags.upload(pb->bc_tags.data(), (int64_t)pb->bc_tags.size(), st);
    pb->d_bc_exprs.upload(pb->bc_exprs.data(), (int64_t)pb->bc_exprs.size(), st);
    RSD_HIP(hipStreamSynchronize(st));
}

static void init_problem(rsd_ctx* ctx, rsd_problem* pb, int32_t dim, const int32_t* p, int32_t nq, int32_t n_bc,
                         const int32_t* bc_tags, const char* const* bc_exprs) {
    RSD_CHECK(dim == 2 || dim == 3, "dim must be 2 or 3");
    for (int d = 1; d < dim; ++d) RSD_CHECK(p[d] == p[0], "the HIP backend needs equal basis orders per direction");
    RSD_CHECK(p[0] >= 1 && p[0] <= 3, "basis order outside [1, 3]");
    RSD_CHECK(nq >= 1 && nq <= 5, "quadrature points outside [1, 5]");
    pb->ctx = ctx;
    pb->dim = dim;
    pb->p = p[0];
    pb->nq = nq;
    pb->npe = dim == 2 ? (p[0] + 1) * (p[0] + 1) : (p[0] + 1) * (p[0] + 1) * (p[0] + 1);
    pb->nfaces = 2 * dim;
    for (int k = 0; k < n_bc; ++k) {
        pb->bc_tags.push_back(bc_tags[k]);
        Expr e = compile_expr(bc_exprs[k] ? bc_exprs[k] : "0");
        if (e.uses_t) pb->bc_time_dependent = true;
        pb->bc_exprs.push_back(e);
    }
}

int rsd_problem_set_assembly(rsd_problem* pb, int32_t mode, int32_t* in_use) {
    return guard([&] {
        RSD_CHECK(pb, "rsd_problem_set_assembly: null problem");
        RSD_CHECK(mode == 0 || mode == 1, "rsd_problem_set_assembly: mode must be 0 (colour scatter) or 1 (row gather)");
        RSD_HIP(hipStreamSynchronize(pb->ctx->stream));
        pb->gather_on = mode == 1;
        if (!pb->gather_on) {  // the element buffers are of no further use
            pb->ebuf.reset();
            pb->fbuf.reset();
        } else if (pb->arith == RSD_ARITH_UNIFORM || pb->arith == RSD_ARITH_FACTORED) {
            arith_reserve(pb);  // so that no later call allocates
        }
        if (in_use) *in_use = pb->gather_on && pb->gat_m > 0 ? 1 : 0;
    });
}

int rsd_problem_set_arithmetic(rsd_problem* pb, int32_t mode) {
    return guard([&] {
        RSD_CHECK(pb, "rsd_problem_set_arithmetic: null problem");
        RSD_CHECK(mode == RSD_ARITH_REFERENCE || mode == RSD_ARITH_CONTRACTED || mode == RSD_ARITH_UNIFORM ||
                      mode == RSD_ARITH_FACTORED,
                  "rsd_problem_set_arithmetic: mode must be RSD_ARITH_REFERENCE, _CONTRACTED, _UNIFORM or _FACTORED");
        pb->arith = mode;
        if (mode == RSD_ARITH_UNIFORM || mode == RSD_ARITH_FACTORED) arith_reserve(pb);
    });
}

int rsd_problem_create_block(rsd_ctx* ctx, int32_t dim, const int64_t* nel, const double* ext, const int32_t* p,
                             int32_t nq, int32_t n_bc, const int32_t* bc_tags, const char* const* bc_exprs,
                             rsd_problem** out) {
    return guard([&] {
        RSD_CHECK(ctx && nel && ext && p && out, "rsd_problem_create_block: null argument");
        auto pb = std::make_unique<rsd_problem>();
        init_problem(ctx, pb.get(), dim, p, nq, n_bc, bc_tags, bc_exprs);
        block_mesh(dim, nel, ext, pb->p, pb->h_xyz, pb->h_ien, pb->h_rng, &pb->n_nodes, &pb->n_elem);
        build_problem(pb.get());
        *out = pb.release();
    });
}

int rsd_block_mesh(int32_t dim, const int64_t* nel, const double* ext, const int32_t* p, double* xyz, int64_t* ien,
                   int32_t* rng) {
    return guard([&] {
        RSD_CHECK(nel && ext && p && xyz && ien && rng, "rsd_block_mesh: null argument");
        RSD_CHECK(dim == 2 || dim == 3, "rsd_block_mesh: dim 2 or 3");
        for (int i = 1; i < dim; ++i) RSD_CHECK(p[i] == p[0], "rsd_block_mesh: equal basis orders per direction");
        std::vector<double> X;
        std::vector<int32_t> I, R;
        int64_t nn = 0, ne = 0;
        block_mesh(dim, nel, ext, p[0], X, I, R, &nn, &ne);
        std::copy(X.begin(), X.end(), xyz);
        for (size_t k = 0; k < I.size(); ++k) ien[k] = I[k];
        std::copy(R.begin(), R.end(), rng);
    });
}

int rsd_problem_create(rsd_ctx* ctx, int32_t dim, const int32_t* p, int32_t nq, int64_t n_nodes, int64_t n_elem,
                       const double* xyz, const int64_t* ien, const int32_t* rng, int32_t n_bc, const int32_t* bc_tags,
                       const char* const* bc_exprs, rsd_problem** out) {
    return guard([&] {
        RSD_CHECK(ctx && p && xyz && ien && rng && out, "rsd_problem_create: null argument");
        auto pb = std::make_unique<rsd_problem>();
        init_problem(ctx, pb.get(), dim, p, nq, n_bc, bc_tags, bc_exprs);
        RSD_CHECK(n_nodes > 0 && n_nodes < INT32_MAX && n_elem > 0, "bad mesh sizes");
        pb->n_nodes = n_nodes;
        pb->n_elem = n_elem;
        pb->h_xyz.assign(xyz, xyz + n_nodes * dim);
        pb->h_ien.resize(n_elem * pb->npe);
        for (int64_t k = 0; k < n_elem * pb->npe; ++k) {
            RSD_CHECK(ien[k] >= 0 && ien[k] < n_nodes, "ien entry out of range");
            pb->h_ien[k] = (int32_t)ien[k];
        }
        pb->h_rng.assign(rng, rng + n_elem * pb->nfaces);
        build_problem(pb.get());
        *out = pb.release();
    });
}

int rsd_problem_eval_expr(rsd_ctx* ctx, rsd_problem* pb, const char* expr, double time, double* d_out) {
    return guard([&] {
        RSD_CHECK(ctx && pb && expr && d_out, "rsd_problem_eval_expr: null argument");
        RSD_CHECK(ctx == pb->ctx, "problem belongs to another context");
        const Expr e = compile_expr(expr);
        if (pb->n_nodes)
            hipLaunchKernelGGL(k_eval_free, dim3((unsigned)nblocks(pb->n_nodes)), dim3(kBlock), 0, ctx->stream,
                               pb->n_nodes, pb->dim, pb->xyz.p, pb->t2a.p, e, time, d_out);
        RSD_HIP(hipGetLastError());
    });
}

int rsd_pointer_is_device(const void* p, int32_t* out) {
    return guard([&] {
        RSD_CHECK(out, "rsd_pointer_is_device: null argument");
        *out = 0;
        if (!p) return;
        hipPointerAttribute_t a{};
        const hipError_t e = hipPointerGetAttributes(&a, p);
        (void)hipGetLastError();  // a plain host pointer is not an error of the caller's
        *out = (e == hipSuccess && a.type == hipMemoryTypeDevice) ? 1 : 0;
    });
}

struct rsd_expr {
    rsd::Expr e;
};

int rsd_expr_compile(const char* text, rsd_expr** out) {
    return guard([&] {
        RSD_CHECK(text && out, "rsd_expr_compile: null argument");
        auto h = std::make_unique<rsd_expr>();
        h->e = compile_expr(text);
        *out = h.release();
    });
}

int rsd_expr_eval(const rsd_expr* e, double t, double x, double y, double z, double T, double* out) {
    return guard([&] {
        RSD_CHECK(e && out, "rsd_expr_eval: null argument");
        *out = eval_expr_host(e->e, t, x, y, z, T);
    });
}

int rsd_expr_uses(const rsd_expr* e, int32_t* uses_t, int32_t* uses_T) {
    return guard([&] {
        RSD_CHECK(e, "rsd_expr_uses: null argument");
        if (uses_t) *uses_t = e->e.uses_t;
        if (uses_T) *uses_T = e->e.uses_T;
    });
}

int rsd_expr_destroy(rsd_expr* e) {
    return guard([&] { delete e; });
}

int rsd_problem_set_dirichlet_values(rsd_ctx* ctx, rsd_problem* pb, double time, const double* values) {
    return guard([&] {
        RSD_CHECK(ctx && pb, "rsd_problem_set_dirichlet_values: null argument");
        RSD_CHECK(ctx == pb->ctx, "problem belongs to another context");
        if (!values) {  // back to the expressions
            pb->gval_external = false;
            pb->gval_valid = false;
            return;
        }
        const int64_t nc = pb->n_nodes - pb->n_free;
        if (pb->cnodes.n != nc) {
            std::vector<int32_t> cn;
            cn.reserve((size_t)nc);
            for (int64_t i = 0; i < pb->n_nodes; ++i)
                if (pb->h_t2a[i] < 0) cn.push_back((int32_t)i);
            RSD_CHECK((int64_t)cn.size() == nc, "rsd_problem_set_dirichlet_values: one DOF per node expected");
            pb->cnodes.upload(cn.data(), nc, ctx->stream);
            pb->cvals.alloc(nc);
            RSD_HIP(hipStreamSynchronize(ctx->stream));
        }
        if (!pb->gval_external) launch_fill(ctx, pb->n_nodes, pb->gval.p, 0.0);
        if (nc) {
            RSD_HIP(hipMemcpyAsync(pb->cvals.p, values, sizeof(double) * (size_t)nc, hipMemcpyHostToDevice, ctx->stream));
            hipLaunchKernelGGL(k_scatter_gval, dim3((unsigned)nblocks(nc)), dim3(kBlock), 0, ctx->stream, nc, pb->cnodes.p,
                               pb->cvals.p, pb->gval.p);
            RSD_HIP(hipGetLastError());
            RSD_HIP(hipStreamSynchronize(ctx->stream));  // values is the caller's host memory
        }
        pb->gval_external = true;
        pb->gval_valid = true;
        pb->gval_time = time;
    });
}

int rsd_problem_set_quadrature(rsd_ctx* ctx, rsd_problem* pb, int32_t nq) {
    return guard([&] {
        RSD_CHECK(ctx && pb, "rsd_problem_set_quadrature: null argument");
        RSD_CHECK(ctx == pb->ctx, "problem belongs to another context");
        RSD_CHECK(nq >= 1 && nq <= 5, "quadrature points outside [1, 5]");
        if (nq == pb->nq) return;
        pb->nq = nq;
        build_shape_table(pb);
        pb->reaction.reset();  // its face points depend on the rule
        RSD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

rsd_problem::~rsd_problem() {
    for (rsd_matrix* m : matrices) m->pb = nullptr;
}

int rsd_problem_destroy(rsd_problem* pb) {
    return guard([&] {
        if (!pb) return;
        (void)hipStreamSynchronize(pb->ctx->stream);
        delete pb;
    });
}

int rsd_problem_sizes(rsd_problem* pb, int64_t* out) {
    return guard([&] {
        RSD_CHECK(pb && out, "null argument");
        out[0] = pb->n_nodes; out[1] = pb->n_elem; out[2] = pb->npe; out[3] = pb->n_free;
        out[4] = pb->nnz; out[5] = pb->n_colours(); out[6] = pb->sell.slots; out[7] = pb->max_row;
    });
}

int rsd_problem_layout(rsd_problem* pb, int64_t* out) {
    return guard([&] {
        RSD_CHECK(pb && out, "null argument");
        out[0] = pb->sell.mode();
        out[1] = pb->sell.n_pat;
        out[2] = pb->sell.has_off16;
        out[3] = pb->sell.slots;
    });
}

int rsd_problem_layout_f32(rsd_problem* pb, int64_t* out) {
    return guard([&] {
        RSD_CHECK(pb && out, "null argument");
        build_half(pb);
        const Sell& S = pb->sell;
        out[0] = S.has_uni ? COL_UNI : S.mode32();
        out[1] = S.has_half ? S.n_half : S.n_pat;
        out[2] = S.has_uni ? S.uni_planes : S.half_width;
        out[3] = S.slots32();
        RSD_HIP(hipStreamSynchronize(pb->ctx->stream));
    });
}

int rsd_problem_padding_f32(rsd_problem* pb, int64_t* lo, int64_t* hi) {
    return guard([&] {
        RSD_CHECK(pb && lo && hi, "null argument");
        build_half(pb);
        *lo = pb->sell.has_uni ? pb->sell.uni_lo : 0;
        *hi = pb->sell.has_uni ? pb->sell.uni_hi : 0;
        RSD_HIP(hipStreamSynchronize(pb->ctx->stream));
    });
}

int rsd_problem_get_mesh(rsd_problem* pb, double* xyz, int64_t* ien, int32_t* rng) {
    return guard([&] {
        RSD_CHECK(pb, "null problem");
        if (xyz) std::copy(pb->h_xyz.begin(), pb->h_xyz.end(), xyz);
        if (ien) std::copy(pb->h_ien.begin(), pb->h_ien.end(), ien);
        if (rng) std::copy(pb->h_rng.begin(), pb->h_rng.end(), rng);
    });
}

int rsd_problem_get_dofs(rsd_problem* pb, int32_t* t2a, int32_t* a2t, int32_t* ctag) {
    return guard([&] {
        RSD_CHECK(pb, "null problem");
        if (t2a) std::copy(pb->h_t2a.begin(), pb->h_t2a.end(), t2a);
        if (a2t) std::copy(pb->h_a2t.begin(), pb->h_a2t.end(), a2t);
        if (ctag) std::copy(pb->h_ctag.begin(), pb->h_ctag.end(), ctag);
    });
}

int rsd_problem_get_pattern(rsd_problem* pb, int64_t* rp, int64_t* ci) {
    return guard([&] {
        RSD_CHECK(pb, "null problem");
        if (rp) std::copy(pb->h_rp.begin(), pb->h_rp.end(), rp);
        if (ci) std::copy(pb->h_ci.begin(), pb->h_ci.end(), ci);
    });
}

int rsd_problem_get_elem_map(rsd_problem* pb, int64_t* map) {
    return guard([&] {
        RSD_CHECK(pb && map, "null argument");
        const int npe = pb->npe;
        parallel_for(pb->n_elem, [&](int64_t lo, int64_t hi) {
            for (int64_t e = lo; e < hi; ++e)
                for (int a = 0; a < npe; ++a) {
                    const int32_t r = pb->h_t2a[pb->h_ien[e * npe + a]];
                    for (int b = 0; b < npe; ++b) {
                        const int32_t c = pb->h_t2a[pb->h_ien[e * npe + b]];
                        int64_t v = -1;
                        if (r >= 0 && c >= 0) {
                            const int32_t* beg = pb->h_ci.data() + pb->h_rp[r];
                            const int32_t* end = pb->h_ci.data() + pb->h_rp[r + 1];
                            const int32_t* it = std::lower_bound(beg, end, c);
                            if (it != end && *it == c) v = pb->h_rp[r] + (it - beg);
                        }
                        map[(e * npe + a) * npe + b] = v;
                    }
                }
        });
    });
}

int rsd_problem_get_colours(rsd_problem* pb, int32_t* colour, int64_t* order) {
    return guard([&] {
        RSD_CHECK(pb, "null problem");
        if (colour) std::copy(pb->h_colour.begin(), pb->h_colour.end(), colour);
        if (order) std::copy(pb->h_order.begin(), pb->h_order.end(), order);
    });
}

int rsd_problem_set_conductivity(rsd_problem* pb, double k_scalar, const char* k_expr, const char* dk_expr,
                                 const double* tensor, const double* kappa) {
    return guard([&] {
        RSD_CHECK(pb, "null problem");
        Model& m = pb->model;
        m.k_scalar = k_scalar;
        m.has_k = k_expr != nullptr;
        m.has_dk = dk_expr != nullptr;
        if (k_expr) m.k = compile_expr(k_expr);
        if (dk_expr) m.dk = compile_expr(dk_expr);
        m.aniso = tensor != nullptr;
        const int d = pb->dim;
        for (int i = 0; i < 9; ++i) m.tensor[i] = 0.0;
        for (int i = 0; i < d * d; ++i) m.tensor[i] = tensor ? tensor[i] : ((i % (d + 1)) == 0 ? 1.0 : 0.0);
        if (kappa) pb->kappa.upload(kappa, pb->n_elem, pb->ctx->stream);
        else pb->kappa.reset();
        pb->kappa_ext = nullptr;
        RSD_HIP(hipStreamSynchronize(pb->ctx->stream));
    });
}

int rsd_problem_set_mass(rsd_problem* pb, double rho, double cp) {
    return guard([&] {
        RSD_CHECK(pb, "null problem");
        pb->model.rho = rho;
        pb->model.cp = cp;
    });
}

int rsd_problem_set_kappa_device(rsd_problem* pb, const double* d_kappa) {
    return guard([&] {
        RSD_CHECK(pb, "null problem");
        pb->kappa_ext = d_kappa;
    });
}

// ------------------------------------------------------------------------
// matrices
// ------------------------------------------------------------------------
int rsd_matrix_create(rsd_problem* pb, rsd_matrix** out) {
    return guard([&] {
        RSD_CHECK(pb && out, "null argument");
        auto m = std::make_unique<rsd_matrix>();
        m->pb = pb;
        m->ctx = pb->ctx;
        m->val.alloc(pb->sell.slots);
        if (pb->sell.slots) RSD_HIP(hipMemsetAsync(m->val.p, 0, sizeof(double) * pb->sell.slots, pb->ctx->stream));
        RSD_HIP(hipStreamSynchronize(pb->ctx->stream));
        pb->matrices.insert(m.get());
        *out = m.release();
    });
}

int rsd_matrix_destroy(rsd_matrix* m) {
    return guard([&] {
        if (!m) return;
        if (m->pb) m->pb->matrices.erase(m);
        if (m->ctx) (void)hipStreamSynchronize(m->ctx->stream);
        delete m;
    });
}

int rsd_matrix_zero(rsd_ctx* ctx, rsd_matrix* m) {
    return guard([&] {
        RSD_CHECK(ctx && m, "null argument");
        RSD_LIVE_MATRIX(m);
        if (m->val.n) RSD_HIP(hipMemsetAsync(m->val.p, 0, sizeof(double) * m->val.n, ctx->stream));
    });
}

static int64_t sell_slot(const rsd_problem* pb, int64_t r, int64_t j) {
    return pb->sell.h_slice_ptr[r / kSlice] + j * kSlice + (r % kSlice);
}

int rsd_matrix_get_values(rsd_ctx* ctx, rsd_matrix* m, double* out) {
    return guard([&] {
        RSD_CHECK(ctx && m && out, "null argument");
        RSD_LIVE_MATRIX(m);
        const rsd_problem* pb = m->pb;
        std::vector<double> h(m->val.n);
        RSD_HIP(hipMemcpyAsync(h.data(), m->val.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
        RSD_HIP(hipStreamSynchronize(ctx->stream));
        for (int64_t r = 0; r < pb->n_free; ++r)
            for (int64_t j = 0; j < pb->h_rp[r + 1] - pb->h_rp[r]; ++j) out[pb->h_rp[r] + j] = h[sell_slot(pb, r, j)];
    });
}

int rsd_matrix_store_csr(rsd_ctx* ctx, rsd_matrix* m, double* d_csr) {
    return guard([&] {
        RSD_CHECK(ctx && m && d_csr, "rsd_matrix_store_csr: null argument");
        RSD_LIVE_MATRIX(m);
        RSD_CHECK(ctx == m->pb->ctx, "matrix belongs to another context");
        launch_sell_csr(m, d_csr, true);
    });
}

int rsd_matrix_set_values(rsd_ctx* ctx, rsd_matrix* m, const double* in) {
    return guard([&] {
        RSD_CHECK(ctx && m && in, "null argument");
        RSD_LIVE_MATRIX(m);
        const rsd_problem* pb = m->pb;
        std::vector<double> h(m->val.n, 0.0);
        for (int64_t r = 0; r < pb->n_free; ++r)
            for (int64_t j = 0; j < pb->h_rp[r + 1] - pb->h_rp[r]; ++j) h[sell_slot(pb, r, j)] = in[pb->h_rp[r] + j];
        RSD_HIP(hipMemcpyAsync(m->val.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, ctx->stream));
        RSD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): ten rounds of the
// 4x32 bijection with multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl key
// increments 0x9E3779B9 / 0xBB67AE85.
static void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int round = 0; round < 10; ++round) {
        const uint64_t a = (uint64_t)0xD2511F53u * c[0], b = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(b >> 32) ^ c[1] ^ k0, n2 = (uint32_t)(a >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)b;
        c[3] = (uint32_t)a;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

static void philox_pair(uint64_t seed, uint32_t stream, int64_t i, double& u1, double& u2) {
    uint32_t c[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), stream, 0u};
    philox4x32(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    auto u53 = [](uint32_t hi, uint32_t lo) {
        const uint64_t v = (((uint64_t)hi << 32) | lo) >> 11;
        return ((double)v + 0.5) * (1.0 / 9007199254740992.0);
    };
    u1 = u53(c[0], c[1]);
    u2 = u53(c[2], c[3]);
}

int rsd_philox_normal(uint64_t seed, uint32_t stream, int64_t n, double* out) {
    return guard([&] {
        RSD_CHECK(n >= 0 && (n == 0 || out), "rsd_philox_normal: bad argument");
        parallel_for(n, [&](int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; ++i) {
                double u1, u2;
                philox_pair(seed, stream, i, u1, u2);
                out[i] = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
            }
        });
    });
}

int rsd_philox_uniform(uint64_t seed, uint32_t stream, int64_t n, double* out) {
    return guard([&] {
        RSD_CHECK(n >= 0 && (n == 0 || out), "rsd_philox_uniform: bad argument");
        parallel_for(n, [&](int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; ++i) {
                double u1, u2;
                philox_pair(seed, stream, i, u1, u2);
                out[i] = u1;
            }
        });
    });
}

}  // extern "C"
