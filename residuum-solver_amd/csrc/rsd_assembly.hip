// rsd_assembly.hip -- element kernels for Q1/Q2 quads and Q1 hexes.
//
// Launched colour by colour: elements of one colour share no node, so the
// scatter through the element->slot map is a plain read-modify-write and the
// summation order per matrix / vector entry is the colour order.  The element math follows the reference operand by
// operand (EvalQuadraturePointVolume.hpp:64-85, JacobianTransform.tpp:3-206,
// Lagrange1D.tpp:4-51, the heat-equation forms) with every product and sum
// rounded separately, so a CPU loop visiting the elements in the same
// colour order produces bitwise identical results.
#include "rsd_assembly.hpp"
#include "rsd_element.hpp"
#include "rsd_linalg.hpp"

#include <algorithm>

namespace rsd {

// Matrix assembly (DIAG = false): thread a scatters row a of Ke through the
// element->slot map.  DIAG = true (Assembler::assembleDiagonal,
// Assembler.tpp:142-197): thread a forms only Ke[a][a], the same chain of
// products and sums as in the matrix, and adds it to diag[row(a)].  MV =
// true: thread a also forms entry a of the unit mass form's element vector
// on the full gather of A.mvU, with k_asm_vector's chain of products and
// sums, and adds it to A.mvF[row(a)] (the same colour order, so bitwise the
// separate vector assembly).  EB = true: the element pass of the row gather
// assembly, over every element in index order: thread a writes row a of Ke
// to A.ebuf[(e * NPE + a) * NPE + b] (and the mass entry to A.fbuf[e * NPE
// + a]) instead of scattering it.
//
// FPAIR = true: the Backward Euler form pair (unscaled isotropic diffusion,
// then the scaled mass form) on a state-free model, straight-line: per
// point and column b, Ke[b] += (sum_i da_i (k db_i)) meas w, then Ke[b] +=
// coef ((rho cp N_a N_b) meas w) -- each entry's chain of the generic loop,
// except that the integrand's sum starts from its first product instead of
// 0 + first product.  The two differ only in the sign of a zero integrand,
// and Ke[b], a sum started at +0, absorbs that: x + (+0) and x + (-0) are
// both x for x != 0, and both +0 for x = +0 (a sum started at +0 is never -0).
//
// FAST = true (with FPAIR; the problem's contracted arithmetic): per point
// the factors s = meas w, k s and coef (rho cp N_a) s are formed once, and
// Ke[b] = fma(coef rho cp N_a s, N_b, fma(k s, dNdx_a . dNdx_b, Ke[b])) with
// the dot product in fused multiply-adds, the mass vector as
// fma((rho cp N_a) s, sum_b N_b Um_b, Fm): 5 FP64 operations per entry and
// point instead of 20.  Entries differ from the reference order by a few
// ulps of the row's largest entry.
template<int D, int P, int NQ, bool STATE, bool DIAG = false, bool MV = false, bool EB = false, bool FPAIR = false,
         bool FAST = false>
__global__ __launch_bounds__((ECfg<D, P, NQ, STATE>::BLOCK)) void k_asm_matrix(AsmArgs A, Model m, int64_t c0,
                                                                             int64_t c1) {
    static_assert(!FPAIR || (!STATE && !DIAG), "the form pair path is for state-free matrices");
    static_assert(!FAST || FPAIR, "contracted arithmetic: the form pair path only");
    using C = ECfg<D, P, NQ, STATE>;
    constexpr int NPE = C::NPE;
    constexpr int NB = DIAG ? 1 : NPE;  // entries of row a this thread forms
    __shared__ double lds[C::EPB * C::ELEM_DOUBLES];
    const int el = threadIdx.x / NPE, a = threadIdx.x % NPE;
    const int64_t k = c0 + (int64_t)blockIdx.x * C::EPB + el;
    const bool valid = k < c1;
    double* rec = lds + el * C::ELEM_DOUBLES;
    int64_t e = 0;
    int32_t node_a = 0;
    double kap = 1.0;
    const bool has_kap = A.kappa != nullptr;
    double Um[MV ? NPE : 1];
    if (valid) {
        e = EB ? k : A.order[k];
        int32_t nodes[NPE];
        double X[NPE][D];
        load_element<D, P>(A, e, nodes, X);
        node_a = nodes[a];
        if constexpr (MV) gather<NPE>(A, nodes, 1, A.mvU, Um);
        if (has_kap) kap = A.kappa[e];
        double Uv[NPE], Ur[NPE];
        if constexpr (STATE) {
            gather<NPE>(A, nodes, 1, A.U, Uv);
            gather<NPE>(A, nodes, 0, A.Udot, Ur);
        }
        phase1<D, P, NQ, STATE>(A, m, nullptr, 0, e, a, X, Uv, Ur, nullptr, kap, has_kap, rec);
    }
    __syncthreads();
    if (!valid) return;
    double kel = m.k_scalar;
    if (has_kap) kel = kap * kel;
    double Ke[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) Ke[i] = 0.0;
    double Fm = 0.0;
    if constexpr (FAST) {
        const double coef = A.coefs[1];
        const double rc = m.rho * m.cp;
#pragma unroll 1
        for (int q = 0; q < C::NQP; ++q) {
            const double* r = rec + q * C::REC;
            const double* tq = A.ntab + q * (NPE + 1);  // uniform: scalar loads
            const double sw = r[C::R_MEAS] * tq[NPE];
            const double ks = kel * sw;
            const double ma = rc * tq[a] * sw;
            const double cm = coef * ma;
            double da[D];
#pragma unroll
            for (int i = 0; i < D; ++i) da[i] = r[C::R_DN + a * D + i];
            double um = 0.0;
#pragma unroll
            for (int b = 0; b < NPE; ++b) {
                const double* db = r + C::R_DN + b * D;
                double dot = da[D - 1] * db[D - 1];
#pragma unroll
                for (int i = D - 2; i >= 0; --i) dot = __builtin_fma(da[i], db[i], dot);
                Ke[b] = __builtin_fma(cm, tq[b], __builtin_fma(ks, dot, Ke[b]));
                if constexpr (MV) um = __builtin_fma(tq[b], Um[b], um);
            }
            if constexpr (MV) Fm = __builtin_fma(ma, um, Fm);
        }
    } else if constexpr (FPAIR) {
        const double coef = A.coefs[1];
#pragma unroll 1
        for (int q = 0; q < C::NQP; ++q) {
            const double* r = rec + q * C::REC;
            const double* tq = A.ntab + q * (NPE + 1);  // uniform: scalar loads
            const double w = tq[NPE];
            double N[NPE];
#pragma unroll
            for (int b = 0; b < NPE; ++b) N[b] = tq[b];
            const double meas = r[C::R_MEAS];
            const double ra = m.rho * m.cp * tq[a];
            double da[D];
#pragma unroll
            for (int i = 0; i < D; ++i) da[i] = r[C::R_DN + a * D + i];
#pragma unroll
            for (int b = 0; b < NPE; ++b) {
                const double rn = ra * N[b];
                if constexpr (MV) Fm += rn * Um[b] * meas * w;
                const double* db = r + C::R_DN + b * D;
                double it = da[0] * (kel * db[0]);
#pragma unroll
                for (int i = 1; i < D; ++i) it += da[i] * (kel * db[i]);
                Ke[b] += it * meas * w;
                Ke[b] += coef * (rn * meas * w);
            }
        }
    } else
    for (int q = 0; q < C::NQP; ++q) {
        const double* r = rec + q * C::REC;
        double xi[3];
        const double w = point_coords<D, NQ>(q, xi);
        double N[NPE];
        shape_values<D, P>(xi, N);
        const double meas = r[C::R_MEAS];
        if constexpr (MV) {
            const double ra = m.rho * m.cp * N[a];
#pragma unroll
            for (int b = 0; b < NPE; ++b) Fm += ra * N[b] * Um[b] * meas * w;
        }
        const double* da = r + C::R_DN + a * D;
        const double kq = STATE ? r[C::R_K] : kel;
        for (int f = 0; f < A.nf; ++f) {
            const int kind = A.kinds[f] & 0xff;
            const bool scaled = (A.kinds[f] & RSD_FORM_SCALED) != 0;
            const double coef = A.coefs[f];
            if (kind == RSD_FORM_DIFFUSION) {
#pragma unroll
                for (int i = 0; i < NB; ++i) {
                    const int b = DIAG ? a : i;
                    const double c = diff_integrand<D>(m, kq, da, r + C::R_DN + b * D) * meas * w;
                    Ke[i] += scaled ? coef * c : c;
                }
            } else if (kind == RSD_FORM_MASS) {
                const double ra = m.rho * m.cp * N[a];
#pragma unroll
                for (int i = 0; i < NB; ++i) {
                    const int b = DIAG ? a : i;
                    const double c = ra * N[b] * meas * w;
                    Ke[i] += scaled ? coef * c : c;
                }
            } else if constexpr (STATE) {
                if (kind == RSD_FORM_TDIFFUSION) {
                    const double it = tdiff_integrand<D>(m, r[C::R_DK], da, r + C::R_GT);
#pragma unroll
                    for (int i = 0; i < NB; ++i) {
                        const int b = DIAG ? a : i;
                        const double c = it * N[b] * meas * w;
                        Ke[i] += scaled ? coef * c : c;
                    }
                } else {  // TMASS: rho * dcp/dT * Trate * N_a * N_b, dcp/dT = 0 for a constant cp
                    const double ra = m.rho * 0.0 * r[C::R_TR] * N[a];
#pragma unroll
                    for (int i = 0; i < NB; ++i) {
                        const int b = DIAG ? a : i;
                        const double c = ra * N[b] * meas * w;
                        Ke[i] += scaled ? coef * c : c;
                    }
                }
            }
        }
    }
    if constexpr (EB) {
        double* out = A.ebuf + (e * NPE + a) * NPE;
#pragma unroll
        for (int b = 0; b < NPE; ++b) out[b] = Ke[b];
        if constexpr (MV) A.fbuf[e * NPE + a] = Fm;
        return;
    }
    if constexpr (MV) {
        const int32_t row = A.t2a[node_a];
        if (row >= 0) A.mvF[row] += Fm;
    }
    if constexpr (DIAG) {
        const int32_t row = A.t2a[node_a];
        if (row >= 0) A.F[row] += Ke[0];
    } else {
        const int32_t* mp = A.map + (e * NPE + a) * NPE;
#pragma unroll
        for (int b = 0; b < NPE; ++b) {
            const int32_t s = mp[b];
            if (s >= 0) A.val[s] += Ke[b];
        }
    }
}

// Row gather pass: one lane per SELL row, one wavefront per slice.  Lane r
// walks the row's elements in colour order and adds element row a's
// entries into a wavefront-private LDS accumulator (acc[k * 64 + lane],
// lane-private, so no barrier) at the row slot position k of each column;
// then writes every slot of its row once, coalesced across the slice, the
// padding slots and padding rows as zeros.  Each entry is 0 plus its
// contributions in colour order: bitwise the colour-by-colour scatter into
// a zeroed value array.  Branch-free: a constrained column, and every
// column of a missing element (past the row's count, read from element row
// 0), goes to dummy slot W + b, so one element's NPE slots are distinct and
// its reads, adds and writes issue as three batches.  MV: the element mass
// vectors summed the same way into mvF[r] (a missing element adds +0, a
// no-op on a sum started at +0).  MJ bounds the elements per row (the host
// picks MJ >= gat_m).
// AX (with the mass vector): also axo[r] = sum_j value_j x[col_j] from 0
// in CSR order, bitwise row_dot_f64 on the values just written (AX 1: the
// row's shared offset pattern, 2: absolute columns).
template<int NPE, int MJ, bool MV, int AX = 0>
__global__ __launch_bounds__(kSlice) void k_gather_rows(const int32_t* __restrict__ gel,
                                                        const uint32_t* __restrict__ gk, int32_t M,
                                                        const double* __restrict__ eb, const double* __restrict__ fb,
                                                        const int64_t* __restrict__ slice_ptr, int64_t n_rows,
                                                        double* __restrict__ val, double* __restrict__ mvF, Cols C,
                                                        const double* __restrict__ axx, double* __restrict__ axo) {
    constexpr int K4 = (NPE + 3) / 4;
    extern __shared__ double acc[];  // [(W + NPE) * 64]: the row slots, then the dummy slots
    const int lane = threadIdx.x;
    const int64_t s = blockIdx.x;
    const int64_t base = slice_ptr[s];
    const int W = (int)((slice_ptr[s + 1] - base) / kSlice);
    for (int k = 0; k < W; ++k) acc[k * kSlice + lane] = 0.0;
    const int32_t* ge = gel + s * M * kSlice + lane;
    const uint32_t* gw = gk + s * M * K4 * kSlice + lane;
    int32_t q[MJ];
    uint32_t kw[MJ][K4];
#pragma unroll
    for (int j = 0; j < MJ; ++j) {
        q[j] = j < M ? ge[j * kSlice] : -1;
#pragma unroll
        for (int t = 0; t < K4; ++t) kw[j][t] = j < M ? gw[(j * K4 + t) * kSlice] : 0xffffffffu;
    }
    double v[MJ][NPE];
    double fv[MJ];
#pragma unroll
    for (int j = 0; j < MJ; ++j) {
        const int64_t row = q[j] < 0 ? 0 : q[j];
        const double* er = eb + row * NPE;
#pragma unroll
        for (int b = 0; b < NPE; ++b) v[j][b] = er[b];
        if constexpr (MV) fv[j] = fb[row];
    }
    double fm = 0.0;
#pragma unroll
    for (int j = 0; j < MJ; ++j) {
        int at[NPE];
        double t[NPE];
#pragma unroll
        for (int b = 0; b < NPE; ++b) {
            const uint32_t k = (kw[j][b / 4] >> (8 * (b % 4))) & 0xffu;
            at[b] = ((k == 0xffu) ? W + b : (int)k) * kSlice + lane;
        }
#pragma unroll
        for (int b = 0; b < NPE; ++b) t[b] = acc[at[b]];
#pragma unroll
        for (int b = 0; b < NPE; ++b) t[b] += v[j][b];
#pragma unroll
        for (int b = 0; b < NPE; ++b) acc[at[b]] = t[b];
        if constexpr (MV) fm += q[j] < 0 ? 0.0 : fv[j];
    }
    for (int k = 0; k < W; ++k) val[base + k * kSlice + lane] = acc[k * kSlice + lane];
    if constexpr (MV) {
        const int64_t r = s * kSlice + lane;
        if (r < n_rows) mvF[r] = fm;
    }
    if constexpr (AX != 0) {
        const int64_t r = s * kSlice + lane;
        if (r < n_rows) {
            double sum = 0.0;
            if constexpr (AX == 1) {
                sum = pat_rows<double>(r, C, [&](const int32_t* po, int len) {
                    double t = 0.0;
                    for (int j = 0; j < len; ++j) t += acc[j * kSlice + lane] * axx[(int)r + po[j]];
                    return t;
                });
            } else {
                const int len = C.rp[r + 1] - C.rp[r];
                for (int j = 0; j < len; ++j) sum += acc[j * kSlice + lane] * axx[C.col[base + (int64_t)j * kSlice + lane]];
            }
            axo[r] = sum;
        }
    }
}

// vector assembly, load vector and essential lift share one kernel:
//   mode 0: forms applied to a selected state buffer (assembleVector)
//   mode 1: load vector (SourceForm / NodalSourceForm)
//   mode 2: essential lift, F -= forms(constrained values)
// EBV (the lift by row gather, mode 2): over every border element at once,
// thread a writes its scatter term -1.0 * Fe to A.ebuf[k * NPE + a] (k the
// border position) instead of adding it to F.
// FPAIR (mode 0, state-free isotropic model, the Backward Euler form pair):
// straight-line per point -- Fe += ((sum_i da_i (k db_i)) Ue_b) meas w over
// b, then buf = sum_b (((rho cp N_a) N_b) Ue_b) meas w and Fe += coef buf --
// the generic loop's chains, the integrand's sum started from its first
// product (a sign of zero at most, which Fe, a sum started at +0, absorbs),
// shape values from the Gauss point table.
// FAST (with FPAIR; contracted arithmetic): per point the operand's
// gradient g = sum_b dNdx_b Ue_b and value u = sum_b N_b Ue_b in fused
// multiply-adds, then Fe = fma(coef rho cp N_a meas w, u, fma(k meas w,
// dNdx_a . g, Fe)): about 40 FP64 operations per point instead of 140.
template<int D, int P, int NQ, bool STATE, bool GRAD, int MODE, bool EBV = false, bool FPAIR = false, bool FAST = false>
__global__ __launch_bounds__((ECfg<D, P, NQ, STATE, GRAD>::BLOCK)) void k_asm_vector(AsmArgs A, Model m, Expr src,
                                                                                   int64_t c0, int64_t c1) {
    static_assert(!EBV || MODE == 2, "element vectors to a buffer: the lift only");
    static_assert(!FPAIR || (MODE == 0 && !STATE && GRAD), "the form pair path: operand / state products");
    static_assert(!FAST || FPAIR, "contracted arithmetic: the form pair path only");
    using C = ECfg<D, P, NQ, STATE, GRAD>;
    constexpr int mode = MODE;
    if (A.stop && *A.stop) return;
    constexpr int NPE = C::NPE;
    __shared__ double lds[C::EPB * C::ELEM_DOUBLES];
    const int el = threadIdx.x / NPE, a = threadIdx.x % NPE;
    const int64_t k = c0 + (int64_t)blockIdx.x * C::EPB + el;
    const bool valid = k < c1;
    double* rec = lds + el * C::ELEM_DOUBLES;
    int64_t e = 0;
    double kap = 1.0;
    const bool has_kap = A.kappa != nullptr;
    int32_t nodes[NPE];
    double Ue[NPE];
    if (valid) {
        e = A.order[k];
        double X[NPE][D];
        load_element<D, P>(A, e, nodes, X);
        if (has_kap) kap = A.kappa[e];
        double Uv[NPE], Ur[NPE];
        if (mode == 0) {
            gather<NPE>(A, nodes, 1, A.U, Uv);
            gather<NPE>(A, nodes, 0, A.Udot, Ur);
            if (A.selector == RSD_SEL_VALUE) {
#pragma unroll
                for (int b = 0; b < NPE; ++b) Ue[b] = Uv[b];
            } else if (A.selector == RSD_SEL_RATE) {
#pragma unroll
                for (int b = 0; b < NPE; ++b) Ue[b] = Ur[b];
            } else {
                gather<NPE>(A, nodes, 0, A.operand, Ue);
            }
        } else if (mode == 1) {
            if (A.nodal) {
#pragma unroll
                for (int b = 0; b < NPE; ++b) Ue[b] = A.nodal[nodes[b]];
            }
        } else {
            gather<NPE>(A, nodes, 2, nullptr, Ue);
            if constexpr (STATE) {
#pragma unroll
                for (int b = 0; b < NPE; ++b) { Uv[b] = 0.0; Ur[b] = 0.0; }
            }
        }
        const int load = mode == 1 ? (A.nodal ? 2 : 1) : 0;
        phase1<D, P, NQ, STATE, GRAD>(A, m, &src, load, e, a, X, Uv, Ur, Ue, kap, has_kap, rec);
    }
    __syncthreads();
    if (!valid) return;
    double kel = m.k_scalar;
    if (has_kap) kel = kap * kel;
    double Fe = 0.0;
    if constexpr (FAST) {
        const double coef = A.coefs[1];
        const double rc = m.rho * m.cp;
#pragma unroll 1
        for (int q = 0; q < C::NQP; ++q) {
            const double* r = rec + q * C::REC;
            const double* tq = A.ntab + q * (NPE + 1);  // uniform: scalar loads
            const double sw = r[C::R_MEAS] * tq[NPE];
            double g[D];
#pragma unroll
            for (int i = 0; i < D; ++i) g[i] = 0.0;
            double u = 0.0;
#pragma unroll
            for (int b = 0; b < NPE; ++b) {
                const double* db = r + C::R_DN + b * D;
#pragma unroll
                for (int i = 0; i < D; ++i) g[i] = __builtin_fma(db[i], Ue[b], g[i]);
                u = __builtin_fma(tq[b], Ue[b], u);
            }
            const double* da = r + C::R_DN + a * D;
            double dg = da[D - 1] * g[D - 1];
#pragma unroll
            for (int i = D - 2; i >= 0; --i) dg = __builtin_fma(da[i], g[i], dg);
            Fe = __builtin_fma(coef * (rc * tq[a] * sw), u, __builtin_fma(kel * sw, dg, Fe));
        }
    } else if constexpr (FPAIR) {
        const double coef = A.coefs[1];
#pragma unroll 1
        for (int q = 0; q < C::NQP; ++q) {
            const double* r = rec + q * C::REC;
            const double* tq = A.ntab + q * (NPE + 1);  // uniform: scalar loads
            const double w = tq[NPE];
            const double meas = r[C::R_MEAS];
            double da[D];
#pragma unroll
            for (int i = 0; i < D; ++i) da[i] = r[C::R_DN + a * D + i];
#pragma unroll
            for (int b = 0; b < NPE; ++b) {
                const double* db = r + C::R_DN + b * D;
                double it = da[0] * (kel * db[0]);
#pragma unroll
                for (int i = 1; i < D; ++i) it += da[i] * (kel * db[i]);
                Fe += it * Ue[b] * meas * w;
            }
            const double ra = m.rho * m.cp * tq[a];
            double buf = 0.0;
#pragma unroll
            for (int b = 0; b < NPE; ++b) buf += ra * tq[b] * Ue[b] * meas * w;
            Fe += coef * buf;
        }
    } else
    for (int q = 0; q < C::NQP; ++q) {
        const double* r = rec + q * C::REC;
        double xi[3];
        const double w = point_coords<D, NQ>(q, xi);
        double N[NPE];
        shape_values<D, P>(xi, N);
        const double meas = r[C::R_MEAS];
        if (mode == 1) {
            Fe += r[C::R_SRC] * N[a] * meas * w;
            continue;
        }
        const double* da = r + C::R_DN + a * D;
        const double kq = STATE ? r[C::R_K] : kel;
        for (int f = 0; f < A.nf; ++f) {
            const int kind = A.kinds[f] & 0xff;
            const bool scaled = (A.kinds[f] & RSD_FORM_SCALED) != 0;
            const double coef = A.coefs[f];
            double buf = 0.0;
            double& acc = scaled ? buf : Fe;
            if (kind == RSD_FORM_DIFFUSION) {
                if constexpr (GRAD) {
#pragma unroll
                    for (int b = 0; b < NPE; ++b)
                        acc += diff_integrand<D>(m, kq, da, r + C::R_DN + b * D) * Ue[b] * meas * w;
                }
            } else if (kind == RSD_FORM_MASS) {
                const double ra = m.rho * m.cp * N[a];
#pragma unroll
                for (int b = 0; b < NPE; ++b) acc += ra * N[b] * Ue[b] * meas * w;
            } else if constexpr (STATE) {
                if (kind == RSD_FORM_TDIFFUSION) {
                    const double it = tdiff_integrand<D>(m, r[C::R_DK], da, r + C::R_GT);
#pragma unroll
                    for (int b = 0; b < NPE; ++b) acc += it * N[b] * Ue[b] * meas * w;
                } else {
                    const double ra = m.rho * 0.0 * r[C::R_TR] * N[a];
#pragma unroll
                    for (int b = 0; b < NPE; ++b) acc += ra * N[b] * Ue[b] * meas * w;
                }
            }
            if (scaled) Fe += coef * buf;
        }
    }
    if constexpr (EBV) {
        A.ebuf[k * NPE + a] = -1.0 * Fe;
        return;
    }
    const int32_t row = A.t2a[nodes[a]];
    if (row >= 0) A.F[row] += mode == 2 ? -1.0 * Fe : Fe;
}

// ------------------------------------------------------------------------
// Uniform geometry (RSD_ARITH_UNIFORM).  On a mesh whose elements are all
// translates of element 0 (a block mesh), the Backward Euler form pair's
// element matrix is kappa_e Kh + coef Mh, with Kh (diffusion at unit
// conductivity, the model's scalar and tensor kept) and Mh (mass) formed
// once on element 0 in the reference's order.  uh holds Kh then Mh, each
// [npe * npe] row-major.  The values differ from the per-element ones by
// the rounding of each element's own coordinates, about 1e-14 relative.
// ------------------------------------------------------------------------

// the matrix row pass with no element pass: k_gather_rows with element row
// a of element e formed in place, fma(kappa_e, Kh[a][b], coef Mh[a][b])
template<int NPE, int MJ, bool MV, int AX = 0>
__global__ __launch_bounds__(kSlice) void k_uniform_rows(const int32_t* __restrict__ gel,
                                                         const uint32_t* __restrict__ gk, int32_t M,
                                                         const double* __restrict__ uh,
                                                         const double* __restrict__ kappa, double coef,
                                                         const double* __restrict__ fb, int64_t n_elem,
                                                         const int64_t* __restrict__ slice_ptr, int64_t n_rows,
                                                         double* __restrict__ val, double* __restrict__ mvF, Cols C,
                                                         const double* __restrict__ axx, double* __restrict__ axo) {
    constexpr int K4 = (NPE + 3) / 4;
    extern __shared__ double acc[];  // [(W + NPE) * 64]: the row slots, then the dummy slots
    const int lane = threadIdx.x;
    const int64_t s = blockIdx.x;
    const int64_t base = slice_ptr[s];
    const int W = (int)((slice_ptr[s + 1] - base) / kSlice);
    for (int k = 0; k < W; ++k) acc[k * kSlice + lane] = 0.0;
    const int32_t* ge = gel + s * M * kSlice + lane;
    const uint32_t* gw = gk + s * M * K4 * kSlice + lane;
    int32_t q[MJ];
    uint32_t kw[MJ][K4];
#pragma unroll
    for (int j = 0; j < MJ; ++j) {
        q[j] = j < M ? ge[j * kSlice] : -1;
#pragma unroll
        for (int t = 0; t < K4; ++t) kw[j][t] = j < M ? gw[(j * K4 + t) * kSlice] : 0xffffffffu;
    }
    double kap[MJ];
    double fv[MJ];
#pragma unroll
    for (int j = 0; j < MJ; ++j) {
        const int64_t row = q[j] < 0 ? 0 : q[j];
        kap[j] = kappa ? kappa[row / NPE] : 1.0;
        if constexpr (MV) fv[j] = fb[(row % NPE) * n_elem + row / NPE];  // by local node
    }
    double fm = 0.0;
#pragma unroll
    for (int j = 0; j < MJ; ++j) {
        const int a = (q[j] < 0 ? 0 : q[j]) % NPE;
        const double* kh = uh + a * NPE;
        const double* mh = uh + NPE * NPE + a * NPE;
        int at[NPE];
        double t[NPE];
#pragma unroll
        for (int b = 0; b < NPE; ++b) {
            const uint32_t k = (kw[j][b / 4] >> (8 * (b % 4))) & 0xffu;
            at[b] = ((k == 0xffu) ? W + b : (int)k) * kSlice + lane;
        }
#pragma unroll
        for (int b = 0; b < NPE; ++b) t[b] = acc[at[b]];
#pragma unroll
        for (int b = 0; b < NPE; ++b) t[b] += __builtin_fma(kap[j], kh[b], coef * mh[b]);
#pragma unroll
        for (int b = 0; b < NPE; ++b) acc[at[b]] = t[b];
        if constexpr (MV) fm += q[j] < 0 ? 0.0 : fv[j];
    }
    for (int k = 0; k < W; ++k) val[base + k * kSlice + lane] = acc[k * kSlice + lane];
    if constexpr (MV) {
        const int64_t r = s * kSlice + lane;
        if (r < n_rows) mvF[r] = fm;
    }
    if constexpr (AX != 0) {
        const int64_t r = s * kSlice + lane;
        if (r < n_rows) {
            double sum = 0.0;
            if constexpr (AX == 1) {
                sum = pat_rows<double>(r, C, [&](const int32_t* po, int len) {
                    double t2 = 0.0;
                    for (int j = 0; j < len; ++j) t2 += acc[j * kSlice + lane] * axx[(int)r + po[j]];
                    return t2;
                });
            } else {
                const int len = C.rp[r + 1] - C.rp[r];
                for (int j = 0; j < len; ++j) sum += acc[j * kSlice + lane] * axx[C.col[base + (int64_t)j * kSlice + lane]];
            }
            axo[r] = sum;
        }
    }
}

// element vectors of the uniform pair, one thread per element, into
// vb[a * n_elem + e] (by local node, so that both the stores here and the
// row passes' loads are coalesced): MODE 0 the operand product (kappa_e Kh +
// coef Mh) Ue with the vector kernel's gathers (selector), MODE 1 the unit
// mass vector Mh Um of the full gather of A.mvU
template<int NPE, int MODE>
__global__ __launch_bounds__(kBlock) void k_uniform_elems(int64_t n_elem, AsmArgs A, const double* __restrict__ uh,
                                                          double* __restrict__ vb) {
    if (MODE == 0 && A.stop && *A.stop) return;
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= n_elem) return;
    int32_t nodes[NPE];
#pragma unroll
    for (int a = 0; a < NPE; ++a) nodes[a] = A.ien[e * NPE + a];
    double Ue[NPE];
    if constexpr (MODE == 1) {
        gather<NPE>(A, nodes, 1, A.mvU, Ue);
    } else {
        if (A.selector == RSD_SEL_VALUE) gather<NPE>(A, nodes, 1, A.U, Ue);
        else if (A.selector == RSD_SEL_RATE) gather<NPE>(A, nodes, 0, A.Udot, Ue);
        else gather<NPE>(A, nodes, 0, A.operand, Ue);
    }
    const double kap = (MODE == 0 && A.kappa) ? A.kappa[e] : 1.0;
    const double coef = A.coefs[1];
#pragma unroll
    for (int a = 0; a < NPE; ++a) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < NPE; ++b) {
            const double mh = uh[NPE * NPE + a * NPE + b];
            const double v = MODE == 1 ? mh : __builtin_fma(kap, uh[a * NPE + b], coef * mh);
            s = __builtin_fma(v, Ue[b], s);
        }
        vb[a * n_elem + e] = s;
    }
}

// out[r] = the row's element terms in colour order, from 0 (one lane per
// row, the row gather tables; vb by local node, vb[a * n_elem + e])
template<int NPE>
__global__ __launch_bounds__(kSlice) void k_sum_rows(const int32_t* __restrict__ gel, int32_t M, int64_t n_elem,
                                                     const double* __restrict__ vb, int64_t n_rows,
                                                     double* __restrict__ out, const int32_t* stop) {
    if (stop && *stop) return;
    const int lane = threadIdx.x;
    const int64_t s = blockIdx.x, r = s * kSlice + lane;
    if (r >= n_rows) return;
    const int32_t* ge = gel + s * M * kSlice + lane;
    double sum = 0.0;
    for (int j = 0; j < M; ++j) {
        const int32_t q = ge[j * kSlice];
        if (q >= 0) sum += vb[(int64_t)(q % NPE) * n_elem + q / NPE];
    }
    out[r] = sum;
}

// ------------------------------------------------------------------------
// Factored operator product (RSD_ARITH_FACTORED): the Backward Euler form
// pair applied to an operand on Q1 elements without forming the element
// matrix.  A Q1 field f = sum_a f_a N_a equals sum_m c_m phi_m, phi_m the
// monomial prod_d xi_d^(bit d of m), with c = T f and T the tensor product
// over the directions of the 1D change of basis (f0, f1) -> ((f0 + f1) / 2,
// (f1 - f0) / 2): the sum factorisation of the Q1 basis.  Per point, the
// Jacobian, the operand's value and its reference gradient come from the
// coefficients (d/dxi_d phi_m = phi_{m - e_d} when bit d of m is set);
// the diffusion flux is f = (k w / |det J|) M^T (M grad_xi u) with M the
// matrix the reference's gradient transform applies times det J
// (JacobianTransform.tpp:71-87 with invertMatrix :159-186: in 3D the
// cofactor matrix, in 2D the adjugate), and the mass term
// s = coef rho cp w |det J| u, where the reference's measure sqrt(det J^T J)
// is |det J|.  The test functions accumulate in coefficient space,
// Z_m = sum_q (phi_m s + sum_d d/dxi_d phi_m f_d), and the element vector is
// T^T Z.  About 125 FP64 operations per point for the whole element vector,
// against 20 per entry and point for the element matrix.
// ------------------------------------------------------------------------
template<int D>
__device__ __forceinline__ void q1_coefficients(double (&f)[1 << D]) {
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
        for (int a = 0; a < (1 << D); ++a)
            if (!(a & (1 << d))) {
                const double lo = f[a], hi = f[a | (1 << d)];
                f[a] = 0.5 * (lo + hi);
                f[a | (1 << d)] = 0.5 * (hi - lo);
            }
}

// z <- T^T z
template<int D>
__device__ __forceinline__ void q1_coefficients_t(double (&z)[1 << D]) {
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
        for (int a = 0; a < (1 << D); ++a)
            if (!(a & (1 << d))) {
                const double lo = z[a], hi = z[a | (1 << d)];
                z[a] = 0.5 * (lo - hi);
                z[a | (1 << d)] = 0.5 * (lo + hi);
            }
}

// one thread per element, into vb[a * n_elem + e] (by local node, the layout
// k_sum_rows reads); the operand by the vector kernel's gathers (selector)
template<int D, int NQ>
__global__ __launch_bounds__(kBlock) void k_factored_elems(int64_t n_elem, AsmArgs A, Model m,
                                                           double* __restrict__ vb) {
    if (A.stop && *A.stop) return;
    constexpr int NPE = 1 << D;
    constexpr int NQP = D == 2 ? NQ * NQ : NQ * NQ * NQ;
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= n_elem) return;
    int32_t nodes[NPE];
    double X[NPE][D];
    load_element<D, 1>(A, e, nodes, X);
    double cu[NPE];
    if (A.selector == RSD_SEL_VALUE) gather<NPE>(A, nodes, 1, A.U, cu);
    else if (A.selector == RSD_SEL_RATE) gather<NPE>(A, nodes, 0, A.Udot, cu);
    else gather<NPE>(A, nodes, 0, A.operand, cu);
    q1_coefficients<D>(cu);
    double cx[D][NPE];
#pragma unroll
    for (int i = 0; i < D; ++i) {
#pragma unroll
        for (int a = 0; a < NPE; ++a) cx[i][a] = X[a][i];
        q1_coefficients<D>(cx[i]);
    }
    const double kel = A.kappa ? A.kappa[e] * m.k_scalar : m.k_scalar;
    const double cm = A.coefs[1] * (m.rho * m.cp);
    double Z[NPE];
#pragma unroll
    for (int a = 0; a < NPE; ++a) Z[a] = 0.0;
    // one point at a time: unrolling the point loop folds the monomials to
    // constants but takes the 3D kernel from 126 to 168 VGPRs
#pragma unroll 1
    for (int q = 0; q < NQP; ++q) {
        double xi[3];
        const double w = point_coords<D, NQ>(q, xi);
        double phi[NPE];
#pragma unroll
        for (int mm = 0; mm < NPE; ++mm) {
            phi[mm] = 1.0;
#pragma unroll
            for (int d = 0; d < D; ++d)
                if (mm & (1 << d)) phi[mm] *= xi[d];
        }
        double u = 0.0, gu[D], J[D][D];
#pragma unroll
        for (int mm = 0; mm < NPE; ++mm) u = __builtin_fma(cu[mm], phi[mm], u);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            gu[d] = 0.0;
#pragma unroll
            for (int i = 0; i < D; ++i) J[i][d] = 0.0;
#pragma unroll
            for (int mm = 0; mm < NPE; ++mm)
                if (mm & (1 << d)) {
                    const double ph = phi[mm ^ (1 << d)];
                    gu[d] = __builtin_fma(cu[mm], ph, gu[d]);
#pragma unroll
                    for (int i = 0; i < D; ++i) J[i][d] = __builtin_fma(cx[i][mm], ph, J[i][d]);
                }
        }
        double M[D][D], det;
        if constexpr (D == 2) {
            M[0][0] = J[1][1]; M[0][1] = -J[0][1];
            M[1][0] = -J[1][0]; M[1][1] = J[0][0];
            det = __builtin_fma(J[0][0], J[1][1], -(J[0][1] * J[1][0]));
        } else {
            M[0][0] = __builtin_fma(J[1][1], J[2][2], -(J[1][2] * J[2][1]));
            M[0][1] = __builtin_fma(J[1][2], J[2][0], -(J[1][0] * J[2][2]));
            M[0][2] = __builtin_fma(J[1][0], J[2][1], -(J[1][1] * J[2][0]));
            M[1][0] = __builtin_fma(J[0][2], J[2][1], -(J[0][1] * J[2][2]));
            M[1][1] = __builtin_fma(J[0][0], J[2][2], -(J[0][2] * J[2][0]));
            M[1][2] = __builtin_fma(J[0][1], J[2][0], -(J[0][0] * J[2][1]));
            M[2][0] = __builtin_fma(J[0][1], J[1][2], -(J[0][2] * J[1][1]));
            M[2][1] = __builtin_fma(J[0][2], J[1][0], -(J[0][0] * J[1][2]));
            M[2][2] = __builtin_fma(J[0][0], J[1][1], -(J[0][1] * J[1][0]));
            det = __builtin_fma(J[0][0], M[0][0], __builtin_fma(J[0][1], M[0][1], J[0][2] * M[0][2]));
        }
        const double ad = fabs(det);
        const double sk = (kel * w) / ad;
        const double s = (cm * w) * ad * u;
        double v[D], f[D];
#pragma unroll
        for (int i = 0; i < D; ++i) {
            v[i] = M[i][0] * gu[0];
#pragma unroll
            for (int d = 1; d < D; ++d) v[i] = __builtin_fma(M[i][d], gu[d], v[i]);
        }
#pragma unroll
        for (int d = 0; d < D; ++d) {
            double t = M[0][d] * v[0];
#pragma unroll
            for (int i = 1; i < D; ++i) t = __builtin_fma(M[i][d], v[i], t);
            f[d] = sk * t;
        }
#pragma unroll
        for (int mm = 0; mm < NPE; ++mm) {
            double z = __builtin_fma(phi[mm], s, Z[mm]);
#pragma unroll
            for (int d = 0; d < D; ++d)
                if (mm & (1 << d)) z = __builtin_fma(phi[mm ^ (1 << d)], f[d], z);
            Z[mm] = z;
        }
    }
    q1_coefficients_t<D>(Z);
#pragma unroll
    for (int a = 0; a < NPE; ++a) vb[a * n_elem + e] = Z[a];
}

// the lift's row pass: each listed row adds its border elements' terms in
// border (colour) order, F = ((F + t_1) + t_2) + ..., bitwise the scatter
__global__ void k_lift_rows(int64_t nrows, const int32_t* __restrict__ rows, const int32_t* __restrict__ off,
                            const int32_t* __restrict__ idx, const double* __restrict__ terms, double* __restrict__ F) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nrows) return;
    const int32_t r = rows[i];
    double v = F[r];
    for (int32_t j = off[i]; j < off[i + 1]; ++j) v += terms[idx[j]];
    F[r] = v;
}

// adjoint contribution per element (no colouring: one output per element):
// grad[e] += -(kappa_e * lam_e . (Khat_e T_e)), Khat_e the diffusion matrix
// at unit conductivity, T_e the full gather of A.U, lam_e the free gather
// of A.operand.  Thread a forms row a of Khat_e T_e; the element's first
// thread adds the rows in order.
template<int D, int P, int NQ>
__global__ __launch_bounds__((ECfg<D, P, NQ, false>::BLOCK)) void k_adj_grad(AsmArgs A, Model m,
                                                                           const double* __restrict__ kappa,
                                                                           double* __restrict__ grad, int64_t n_elem) {
    using C = ECfg<D, P, NQ, false>;
    constexpr int NPE = C::NPE;
    __shared__ double lds[C::EPB * C::ELEM_DOUBLES];
    __shared__ double rows[C::EPB * NPE];
    const int el = threadIdx.x / NPE, a = threadIdx.x % NPE;
    const int64_t e = (int64_t)blockIdx.x * C::EPB + el;
    const bool valid = e < n_elem;
    double* rec = lds + el * C::ELEM_DOUBLES;
    int32_t nodes[NPE];
    double Te[NPE];
    if (valid) {
        double X[NPE][D];
        load_element<D, P>(A, e, nodes, X);
        gather<NPE>(A, nodes, 1, A.U, Te);
        phase1<D, P, NQ, false>(A, m, nullptr, 0, e, a, X, nullptr, nullptr, nullptr, 1.0, false, rec);
    }
    __syncthreads();
    if (valid) {
        double Ke[NPE];
#pragma unroll
        for (int b = 0; b < NPE; ++b) Ke[b] = 0.0;
        for (int q = 0; q < C::NQP; ++q) {
            const double* r = rec + q * C::REC;
            double xi[3];
            const double w = point_coords<D, NQ>(q, xi);
            const double meas = r[C::R_MEAS];
            const double* da = r + C::R_DN + a * D;
#pragma unroll
            for (int b = 0; b < NPE; ++b) Ke[b] += diff_integrand<D>(m, m.k_scalar, da, r + C::R_DN + b * D) * meas * w;
        }
        double v = 0.0;
#pragma unroll
        for (int b = 0; b < NPE; ++b) v += Ke[b] * Te[b];
        const int32_t ra = A.t2a[nodes[a]];
        const double la = (ra >= 0 && A.operand) ? A.operand[ra] : 0.0;
        rows[el * NPE + a] = la * v;
    }
    __syncthreads();
    if (valid && a == 0) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < NPE; ++b) s += rows[el * NPE + b];
        grad[e] += -(kappa[e] * s);
    }
}

template<int D, int P, int NQ>
static void run_adj(rsd_problem* pb, const AsmArgs& A, const Model& unit, const double* kappa, double* grad) {
    using C = ECfg<D, P, NQ, false>;
    if (pb->n_elem)
        hipLaunchKernelGGL((k_adj_grad<D, P, NQ>), dim3((unsigned)((pb->n_elem + C::EPB - 1) / C::EPB)),
                           dim3(C::BLOCK), 0, pb->ctx->stream, A, unit, kappa, grad, pb->n_elem);
    RSD_HIP(hipGetLastError());
}

// the Backward Euler form pair on a state-free isotropic model
static bool form_pair(const rsd_problem* pb, const AsmArgs& A) {
    return !A.need_state && !pb->model.aniso && A.nf == 2 && A.kinds[0] == RSD_FORM_DIFFUSION &&
           A.kinds[1] == (RSD_FORM_MASS | RSD_FORM_SCALED);
}

template<int D, int P, int NQ, bool STATE, bool DIAG, bool MV = false>
static void run_matrix_t(rsd_problem* pb, const AsmArgs& A) {
    using C = ECfg<D, P, NQ, STATE>;
    hipStream_t st = pb->ctx->stream;
    const bool pair = !STATE && !DIAG && form_pair(pb, A);
    for (int c = 0; c < pb->n_colours(); ++c) {
        const int64_t c0 = pb->colour_off[c], c1 = pb->colour_off[c + 1];
        if (c1 <= c0) continue;
        const dim3 grid((unsigned)((c1 - c0 + C::EPB - 1) / C::EPB));
        if constexpr (!STATE && !DIAG) {
            if (pair) {
                if (pb->arith != RSD_ARITH_REFERENCE)
                    hipLaunchKernelGGL((k_asm_matrix<D, P, NQ, false, false, MV, false, true, true>), grid,
                                       dim3(C::BLOCK), 0, st, A, pb->model, c0, c1);
                else
                    hipLaunchKernelGGL((k_asm_matrix<D, P, NQ, false, false, MV, false, true>), grid, dim3(C::BLOCK),
                                       0, st, A, pb->model, c0, c1);
                continue;
            }
        }
        hipLaunchKernelGGL((k_asm_matrix<D, P, NQ, STATE, DIAG, MV>), grid, dim3(C::BLOCK), 0, st, A, pb->model, c0,
                           c1);
    }
    RSD_HIP(hipGetLastError());
}

template<int D, int P, int NQ>
__global__ void k_shape_table(double* __restrict__ tab) {
    constexpr int NPE = Elem<D, P>::NPE;
    constexpr int NQP = D == 2 ? NQ * NQ : NQ * NQ * NQ;
    const int q = threadIdx.x;
    if (q >= NQP) return;
    double xi[3];
    const double w = point_coords<D, NQ>(q, xi);
    double N[NPE];
    shape_values<D, P>(xi, N);
    for (int b = 0; b < NPE; ++b) tab[q * (NPE + 1) + b] = N[b];
    tab[q * (NPE + 1) + NPE] = w;
}

template<int D, int P, int NQ>
static void run_shape_table(rsd_problem* pb) {
    constexpr int NPE = Elem<D, P>::NPE;
    constexpr int NQP = D == 2 ? NQ * NQ : NQ * NQ * NQ;
    pb->ntab.alloc(NQP * (NPE + 1));
    hipLaunchKernelGGL((k_shape_table<D, P, NQ>), dim3(1), dim3(128), 0, pb->ctx->stream, pb->ntab.p);
    RSD_HIP(hipGetLastError());
}

template<int NPE, int MJ, bool MV>
static void launch_gather_rows(rsd_problem* pb, const AsmArgs& A) {
    const Sell& S = pb->sell;
    const dim3 grid((unsigned)S.n_slices), block(kSlice);
    const size_t lds = (size_t)(pb->max_row + NPE) * kSlice * sizeof(double);
    hipStream_t st = pb->ctx->stream;
    const Cols C = S.cols();
    if (MV && A.axo) {
        if (S.has_pat)
            hipLaunchKernelGGL((k_gather_rows<NPE, MJ, true, 1>), grid, block, lds, st, pb->gat_el.p, pb->gat_k.p,
                               pb->gat_m, A.ebuf, A.fbuf, S.slice_ptr.p, S.n_rows, A.val, A.mvF, C, A.axx, A.axo);
        else
            hipLaunchKernelGGL((k_gather_rows<NPE, MJ, true, 2>), grid, block, lds, st, pb->gat_el.p, pb->gat_k.p,
                               pb->gat_m, A.ebuf, A.fbuf, S.slice_ptr.p, S.n_rows, A.val, A.mvF, C, A.axx, A.axo);
        return;
    }
    hipLaunchKernelGGL((k_gather_rows<NPE, MJ, MV>), grid, block, lds, st, pb->gat_el.p, pb->gat_k.p, pb->gat_m,
                       A.ebuf, A.fbuf, S.slice_ptr.p, S.n_rows, A.val, A.mvF, C, A.axx, A.axo);
}

static bool use_gather(const rsd_problem* pb) { return pb->gather_on && pb->gat_m > 0; }

bool gather_forms_product(const rsd_problem* pb) { return use_gather(pb); }

void gather_reserve(rsd_problem* pb, bool with_mass) {
    if (!use_gather(pb)) return;
    const int64_t ne = pb->n_elem, npe = pb->npe;
    if (pb->ebuf.n < ne * npe * npe) pb->ebuf.alloc(ne * npe * npe);
    if (with_mass && pb->fbuf.n < ne * npe) pb->fbuf.alloc(ne * npe);
}

// every element a translate of element 0 (its node offsets from node 0
// equal element 0's within 1e-10 of element 0's extent): checked once
static bool uniform_geometry(rsd_problem* pb) {
    if (pb->uniform_geom >= 0) return pb->uniform_geom == 1;
    const int npe = pb->npe, D = pb->dim;
    bool ok = pb->n_elem > 0 && (int64_t)pb->h_ien.size() == pb->n_elem * npe &&
              (int64_t)pb->h_xyz.size() == pb->n_nodes * D;
    std::vector<double> ref((size_t)npe * D);
    double ext = 0.0;
    if (ok) {
        const int32_t* n0 = pb->h_ien.data();
        for (int a = 0; a < npe; ++a)
            for (int i = 0; i < D; ++i) {
                ref[a * D + i] = pb->h_xyz[(size_t)n0[a] * D + i] - pb->h_xyz[(size_t)n0[0] * D + i];
                ext = std::max(ext, std::fabs(ref[a * D + i]));
            }
        ok = ext > 0.0;
    }
    const double tol = 1e-10 * ext;
    for (int64_t e = 1; ok && e < pb->n_elem; ++e) {
        const int32_t* ne = pb->h_ien.data() + e * npe;
        for (int a = 1; ok && a < npe; ++a)
            for (int i = 0; i < D; ++i) {
                const double d = pb->h_xyz[(size_t)ne[a] * D + i] - pb->h_xyz[(size_t)ne[0] * D + i];
                if (!(std::fabs(d - ref[a * D + i]) <= tol)) { ok = false; break; }
            }
    }
    pb->uniform_geom = ok ? 1 : 0;
    return ok;
}

static bool use_uniform(rsd_problem* pb, const AsmArgs& A) {
    return pb->arith == RSD_ARITH_UNIFORM && use_gather(pb) && form_pair(pb, A) && uniform_geometry(pb);
}

// the factored operator product: Q1 elements, the row gather's tables
static bool use_factored(const rsd_problem* pb, const AsmArgs& A) {
    return pb->arith == RSD_ARITH_FACTORED && pb->p == 1 && use_gather(pb) && form_pair(pb, A);
}

// Kh (unit conductivity) and Mh of element 0 by the element kernel, in the
// reference's order, into pb->uhat
template<int D, int P, int NQ>
static const double* uniform_tables(rsd_problem* pb) {
    using C = ECfg<D, P, NQ, false>;
    constexpr int NPE = C::NPE;
    if (pb->uhat.n < 2 * NPE * NPE) pb->uhat.alloc(2 * NPE * NPE);
    for (int f = 0; f < 2; ++f) {
        AsmArgs A = base_args(pb);
        A.kappa = nullptr;
        A.nf = 1;
        A.kinds[0] = f == 0 ? RSD_FORM_DIFFUSION : RSD_FORM_MASS;
        A.coefs[0] = 1.0;
        A.ebuf = pb->uhat.p + f * NPE * NPE;
        hipLaunchKernelGGL((k_asm_matrix<D, P, NQ, false, false, false, true>), dim3(1), dim3(C::BLOCK), 0,
                           pb->ctx->stream, A, pb->model, (int64_t)0, (int64_t)1);
    }
    RSD_HIP(hipGetLastError());
    return pb->uhat.p;
}

template<int NPE, int MJ, bool MV>
static void launch_uniform_rows(rsd_problem* pb, const AsmArgs& A, const double* uh) {
    const Sell& S = pb->sell;
    const dim3 grid((unsigned)S.n_slices), block(kSlice);
    const size_t lds = (size_t)(pb->max_row + NPE) * kSlice * sizeof(double);
    hipStream_t st = pb->ctx->stream;
    const Cols C = S.cols();
    const double coef = A.coefs[1];
    if (MV && A.axo) {
        if (S.has_pat)
            hipLaunchKernelGGL((k_uniform_rows<NPE, MJ, true, 1>), grid, block, lds, st, pb->gat_el.p, pb->gat_k.p,
                               pb->gat_m, uh, A.kappa, coef, A.fbuf, pb->n_elem, S.slice_ptr.p, S.n_rows, A.val, A.mvF, C, A.axx,
                               A.axo);
        else
            hipLaunchKernelGGL((k_uniform_rows<NPE, MJ, true, 2>), grid, block, lds, st, pb->gat_el.p, pb->gat_k.p,
                               pb->gat_m, uh, A.kappa, coef, A.fbuf, pb->n_elem, S.slice_ptr.p, S.n_rows, A.val, A.mvF, C, A.axx,
                               A.axo);
        return;
    }
    hipLaunchKernelGGL((k_uniform_rows<NPE, MJ, MV>), grid, block, lds, st, pb->gat_el.p, pb->gat_k.p, pb->gat_m, uh,
                       A.kappa, coef, A.fbuf, pb->n_elem, S.slice_ptr.p, S.n_rows, A.val, A.mvF, C, A.axx, A.axo);
}

// the uniform pair's matrix (and mass vector): no element pass, no element
// buffer; the mass vector's element terms by a light element pass
template<int D, int P, int NQ, bool MV>
static void run_matrix_uniform(rsd_problem* pb, AsmArgs A) {
    constexpr int NPE = Elem<D, P>::NPE;
    const double* uh = uniform_tables<D, P, NQ>(pb);
    const int64_t ne = pb->n_elem;
    if constexpr (MV) {
        if (pb->fbuf.n < ne * NPE) pb->fbuf.alloc(ne * NPE);
        A.fbuf = pb->fbuf.p;
        hipLaunchKernelGGL((k_uniform_elems<NPE, 1>), dim3((unsigned)nblocks(ne)), dim3(kBlock), 0, pb->ctx->stream,
                           ne, A, uh, A.fbuf);
    }
    const int M = pb->gat_m;
    if (M <= 4) launch_uniform_rows<NPE, 4, MV>(pb, A, uh);
    else if (M <= 8) launch_uniform_rows<NPE, 8, MV>(pb, A, uh);
    else launch_uniform_rows<NPE, kGatherMaxElems, MV>(pb, A, uh);
    RSD_HIP(hipGetLastError());
}

// the uniform pair applied to a state or operand (mode 0): element vectors,
// then each row sums its elements' terms in colour order into A.F
template<int D, int P, int NQ>
static void run_vector_uniform(rsd_problem* pb, const AsmArgs& A) {
    constexpr int NPE = Elem<D, P>::NPE;
    const double* uh = uniform_tables<D, P, NQ>(pb);
    const int64_t ne = pb->n_elem;
    if (pb->vbuf.n < ne * NPE) pb->vbuf.alloc(ne * NPE);
    hipStream_t st = pb->ctx->stream;
    hipLaunchKernelGGL((k_uniform_elems<NPE, 0>), dim3((unsigned)nblocks(ne)), dim3(kBlock), 0, st, ne, A, uh,
                       pb->vbuf.p);
    hipLaunchKernelGGL(k_sum_rows<NPE>, dim3((unsigned)pb->sell.n_slices), dim3(kSlice), 0, st, pb->gat_el.p,
                       pb->gat_m, ne, pb->vbuf.p, pb->n_free, A.F, A.stop);
    RSD_HIP(hipGetLastError());
}

// the pair applied to a state or operand in factored form (Q1): element
// vectors, then each row sums its elements' terms in colour order into A.F
template<int D, int P, int NQ>
static void run_vector_factored(rsd_problem* pb, const AsmArgs& A) {
    if constexpr (P == 1) {
        constexpr int NPE = Elem<D, P>::NPE;
        const int64_t ne = pb->n_elem;
        if (pb->vbuf.n < ne * NPE) pb->vbuf.alloc(ne * NPE);
        hipStream_t st = pb->ctx->stream;
        hipLaunchKernelGGL((k_factored_elems<D, NQ>), dim3((unsigned)nblocks(ne)), dim3(kBlock), 0, st, ne, A,
                           pb->model, pb->vbuf.p);
        hipLaunchKernelGGL(k_sum_rows<NPE>, dim3((unsigned)pb->sell.n_slices), dim3(kSlice), 0, st, pb->gat_el.p,
                           pb->gat_m, ne, pb->vbuf.p, pb->n_free, A.F, A.stop);
        RSD_HIP(hipGetLastError());
    } else {
        throw Error(RSD_ERR_INTERNAL, "factored operator product: Q1 elements only");
    }
}

// row gather assembly: the element pass (one launch over all elements),
// then the row pass
template<int D, int P, int NQ, bool STATE, bool MV>
static void run_matrix_gather(rsd_problem* pb, AsmArgs A) {
    using C = ECfg<D, P, NQ, STATE>;
    constexpr int NPE = C::NPE;
    if constexpr (!STATE) {
        if (use_uniform(pb, A)) {
            run_matrix_uniform<D, P, NQ, MV>(pb, A);
            return;
        }
    }
    gather_reserve(pb, MV);
    A.ebuf = pb->ebuf.p;
    A.fbuf = MV ? pb->fbuf.p : nullptr;
    const int64_t ne = pb->n_elem;
    const dim3 grid((unsigned)((ne + C::EPB - 1) / C::EPB));
    bool done = false;
    if constexpr (!STATE) {
        if (form_pair(pb, A)) {
            if (pb->arith != RSD_ARITH_REFERENCE)
                hipLaunchKernelGGL((k_asm_matrix<D, P, NQ, false, false, MV, true, true, true>), grid, dim3(C::BLOCK),
                                   0, pb->ctx->stream, A, pb->model, (int64_t)0, ne);
            else
                hipLaunchKernelGGL((k_asm_matrix<D, P, NQ, false, false, MV, true, true>), grid, dim3(C::BLOCK), 0,
                                   pb->ctx->stream, A, pb->model, (int64_t)0, ne);
            done = true;
        }
    }
    if (!done)
        hipLaunchKernelGGL((k_asm_matrix<D, P, NQ, STATE, false, MV, true>), grid, dim3(C::BLOCK), 0, pb->ctx->stream,
                           A, pb->model, (int64_t)0, ne);
    RSD_HIP(hipGetLastError());
    const int M = pb->gat_m;
    if (M <= 4) launch_gather_rows<NPE, 4, MV>(pb, A);
    else if (M <= 8) launch_gather_rows<NPE, 8, MV>(pb, A);
    else launch_gather_rows<NPE, kGatherMaxElems, MV>(pb, A);
    RSD_HIP(hipGetLastError());
}

// A.F set: the diagonal (assembleDiagonal) into A.F; else the matrix into
// A.val, with the mass vector of A.mvU into A.mvF when A.mvF is set (row
// gather assembly when the problem has its tables, else colour by colour)
template<int D, int P, int NQ>
static void run_matrix(rsd_problem* pb, const AsmArgs& A) {
    const bool diag = A.F != nullptr, mv = A.mvF != nullptr;
    if (!diag && use_gather(pb)) {
        if (A.need_state) {
            if (mv) run_matrix_gather<D, P, NQ, true, true>(pb, A);
            else run_matrix_gather<D, P, NQ, true, false>(pb, A);
        } else {
            if (mv) run_matrix_gather<D, P, NQ, false, true>(pb, A);
            else run_matrix_gather<D, P, NQ, false, false>(pb, A);
        }
        return;
    }
    if (A.need_state) {
        if (diag) run_matrix_t<D, P, NQ, true, true>(pb, A);
        else if (mv) run_matrix_t<D, P, NQ, true, false, true>(pb, A);
        else run_matrix_t<D, P, NQ, true, false>(pb, A);
    } else {
        if (diag) run_matrix_t<D, P, NQ, false, true>(pb, A);
        else if (mv) run_matrix_t<D, P, NQ, false, false, true>(pb, A);
        else run_matrix_t<D, P, NQ, false, false>(pb, A);
    }
}

template<int D, int P, int NQ, bool STATE, bool GRAD, int MODE>
static void run_vector_t(rsd_problem* pb, AsmArgs A, const Expr& src) {
    using C = ECfg<D, P, NQ, STATE, GRAD>;
    hipStream_t st = pb->ctx->stream;
    const bool lift = MODE == 2;
    if (lift) A.order = pb->border.p;
    if constexpr (MODE == 2) {
        if (use_gather(pb) && pb->lift_nrows > 0) {  // two launches instead of one per colour
            const int64_t nbd = pb->border.n;
            A.ebuf = pb->lift_buf.p;
            hipLaunchKernelGGL((k_asm_vector<D, P, NQ, STATE, GRAD, 2, true>),
                               dim3((unsigned)((nbd + C::EPB - 1) / C::EPB)), dim3(C::BLOCK), 0, st, A, pb->model,
                               src, (int64_t)0, nbd);
            hipLaunchKernelGGL(k_lift_rows, dim3((unsigned)nblocks(pb->lift_nrows)), dim3(kBlock), 0, st,
                               pb->lift_nrows, pb->lift_row.p, pb->lift_off.p, pb->lift_idx.p, pb->lift_buf.p, A.F);
            RSD_HIP(hipGetLastError());
            return;
        }
    }
    const std::vector<int64_t>& off = lift ? pb->bcolour_off : pb->colour_off;
    bool pair = false;
    if constexpr (MODE == 0 && !STATE && GRAD) {
        if (use_uniform(pb, A)) {
            run_vector_uniform<D, P, NQ>(pb, A);
            return;
        }
        if (use_factored(pb, A)) {
            run_vector_factored<D, P, NQ>(pb, A);
            return;
        }
        pair = form_pair(pb, A);
    }
    for (int c = 0; c < pb->n_colours(); ++c) {
        const int64_t c0 = off[c], c1 = off[c + 1];
        if (c1 <= c0) continue;
        const dim3 grid((unsigned)((c1 - c0 + C::EPB - 1) / C::EPB));
        if constexpr (MODE == 0 && !STATE && GRAD) {
            if (pair) {
                if (pb->arith != RSD_ARITH_REFERENCE)
                    hipLaunchKernelGGL((k_asm_vector<D, P, NQ, false, true, 0, false, true, true>), grid,
                                       dim3(C::BLOCK), 0, st, A, pb->model, src, c0, c1);
                else
                    hipLaunchKernelGGL((k_asm_vector<D, P, NQ, false, true, 0, false, true>), grid, dim3(C::BLOCK), 0,
                                       st, A, pb->model, src, c0, c1);
                continue;
            }
        }
        hipLaunchKernelGGL((k_asm_vector<D, P, NQ, STATE, GRAD, MODE>), grid, dim3(C::BLOCK), 0, st, A, pb->model,
                           src, c0, c1);
    }
    RSD_HIP(hipGetLastError());
}

template<int D, int P, int NQ>
static void run_vector(rsd_problem* pb, const AsmArgs& A, const Expr& src, int mode) {
    bool grad = false;  // a load vector or mass forms only: no J^-1, no dN/dx
    if (mode != 1)
        for (int f = 0; f < A.nf; ++f) {
            const int k = A.kinds[f] & 0xff;
            if (k == RSD_FORM_DIFFUSION || k == RSD_FORM_TDIFFUSION) grad = true;
        }
    // one kernel per mode, so that the load vector's expression interpreter
    // does not set the register budget of the others
    if (mode == 1) run_vector_t<D, P, NQ, false, false, 1>(pb, A, src);
    else if (mode == 0) {
        if (A.need_state) run_vector_t<D, P, NQ, true, true, 0>(pb, A, src);
        else if (grad) run_vector_t<D, P, NQ, false, true, 0>(pb, A, src);
        else run_vector_t<D, P, NQ, false, false, 0>(pb, A, src);
    } else {
        if (A.need_state) run_vector_t<D, P, NQ, true, true, 2>(pb, A, src);
        else if (grad) run_vector_t<D, P, NQ, false, true, 2>(pb, A, src);
        else run_vector_t<D, P, NQ, false, false, 2>(pb, A, src);
    }
}

#define RSD_DISPATCH(D_, P_, NQ_, CALL)                                                      \
    if (pb->dim == D_ && pb->p == P_ && pb->nq == NQ_) { CALL<D_, P_, NQ_>; return; }

static void dispatch_matrix(rsd_problem* pb, const AsmArgs& A) {
#define RSD_M(D_, P_, NQ_) RSD_DISPATCH(D_, P_, NQ_, run_matrix)
    if (pb->dim == 2 && pb->p == 1 && pb->nq == 1) { run_matrix<2, 1, 1>(pb, A); return; }
    if (pb->dim == 2 && pb->p == 1 && pb->nq == 2) { run_matrix<2, 1, 2>(pb, A); return; }
    if (pb->dim == 2 && pb->p == 1 && pb->nq == 3) { run_matrix<2, 1, 3>(pb, A); return; }
    if (pb->dim == 2 && pb->p == 2 && pb->nq == 2) { run_matrix<2, 2, 2>(pb, A); return; }
    if (pb->dim == 2 && pb->p == 2 && pb->nq == 3) { run_matrix<2, 2, 3>(pb, A); return; }
    if (pb->dim == 2 && pb->p == 2 && pb->nq == 4) { run_matrix<2, 2, 4>(pb, A); return; }
    if (pb->dim == 3 && pb->p == 1 && pb->nq == 1) { run_matrix<3, 1, 1>(pb, A); return; }
    if (pb->dim == 3 && pb->p == 1 && pb->nq == 2) { run_matrix<3, 1, 2>(pb, A); return; }
    if (pb->dim == 3 && pb->p == 1 && pb->nq == 3) { run_matrix<3, 1, 3>(pb, A); return; }
#undef RSD_M
    throw Error(RSD_ERR_ARG, "discretisation not supported by the HIP element kernels (dim/p/nq)");
}

void build_shape_table(rsd_problem* pb) {
    if (pb->dim == 2 && pb->p == 1 && pb->nq == 1) { run_shape_table<2, 1, 1>(pb); return; }
    if (pb->dim == 2 && pb->p == 1 && pb->nq == 2) { run_shape_table<2, 1, 2>(pb); return; }
    if (pb->dim == 2 && pb->p == 1 && pb->nq == 3) { run_shape_table<2, 1, 3>(pb); return; }
    if (pb->dim == 2 && pb->p == 2 && pb->nq == 2) { run_shape_table<2, 2, 2>(pb); return; }
    if (pb->dim == 2 && pb->p == 2 && pb->nq == 3) { run_shape_table<2, 2, 3>(pb); return; }
    if (pb->dim == 2 && pb->p == 2 && pb->nq == 4) { run_shape_table<2, 2, 4>(pb); return; }
    if (pb->dim == 3 && pb->p == 1 && pb->nq == 1) { run_shape_table<3, 1, 1>(pb); return; }
    if (pb->dim == 3 && pb->p == 1 && pb->nq == 2) { run_shape_table<3, 1, 2>(pb); return; }
    if (pb->dim == 3 && pb->p == 1 && pb->nq == 3) { run_shape_table<3, 1, 3>(pb); return; }
    pb->ntab.reset();  // other discretisations have no element kernels
}

static void dispatch_vector(rsd_problem* pb, const AsmArgs& A, const Expr& src, int mode) {
    if (pb->dim == 2 && pb->p == 1 && pb->nq == 1) { run_vector<2, 1, 1>(pb, A, src, mode); return; }
    if (pb->dim == 2 && pb->p == 1 && pb->nq == 2) { run_vector<2, 1, 2>(pb, A, src, mode); return; }
    if (pb->dim == 2 && pb->p == 1 && pb->nq == 3) { run_vector<2, 1, 3>(pb, A, src, mode); return; }
    if (pb->dim == 2 && pb->p == 2 && pb->nq == 2) { run_vector<2, 2, 2>(pb, A, src, mode); return; }
    if (pb->dim == 2 && pb->p == 2 && pb->nq == 3) { run_vector<2, 2, 3>(pb, A, src, mode); return; }
    if (pb->dim == 2 && pb->p == 2 && pb->nq == 4) { run_vector<2, 2, 4>(pb, A, src, mode); return; }
    if (pb->dim == 3 && pb->p == 1 && pb->nq == 1) { run_vector<3, 1, 1>(pb, A, src, mode); return; }
    if (pb->dim == 3 && pb->p == 1 && pb->nq == 2) { run_vector<3, 1, 2>(pb, A, src, mode); return; }
    if (pb->dim == 3 && pb->p == 1 && pb->nq == 3) { run_vector<3, 1, 3>(pb, A, src, mode); return; }
    throw Error(RSD_ERR_ARG, "discretisation not supported by the HIP element kernels (dim/p/nq)");
}

AsmArgs base_args(rsd_problem* pb) {
    AsmArgs A{};
    A.xyz = pb->xyz.p;
    A.ien = pb->ien.p;
    A.t2a = pb->t2a.p;
    A.gval = pb->gval.p;
    A.order = pb->order.p;
    A.map = pb->map_sell.p;
    A.kappa = pb->kappa_ptr();
    A.ntab = pb->ntab.p;
    return A;
}

static void set_forms(AsmArgs& A, int32_t nf, const int32_t* kinds, const double* coefs) {
    RSD_CHECK(nf >= 0 && nf <= 4, "between 0 and 4 forms");
    A.nf = nf;
    for (int f = 0; f < nf; ++f) {
        const int k = kinds[f] & 0xff;
        RSD_CHECK(k >= RSD_FORM_DIFFUSION && k <= RSD_FORM_TMASS, "unknown form kind");
        A.kinds[f] = kinds[f];
        A.coefs[f] = coefs ? coefs[f] : 1.0;
    }
}

static bool forms_need_state(const rsd_problem* pb, const AsmArgs& A) {
    bool need = pb->model.has_k || pb->model.has_dk;
    for (int f = 0; f < A.nf; ++f) {
        const int k = A.kinds[f] & 0xff;
        if (k == RSD_FORM_TDIFFUSION || k == RSD_FORM_TMASS) need = true;
    }
    return need;
}

// the uniform-geometry arithmetic's device buffers (tables, element vectors
// for the mass vector and for operand products), or the factored product's
// element vectors, reserved when the mode is selected so that no later call
// allocates, under stream capture or not
void arith_reserve(rsd_problem* pb) {
    if (pb->arith == RSD_ARITH_FACTORED) {  // the element vectors of the factored product
        if (use_gather(pb) && pb->p == 1 && pb->vbuf.n < pb->n_elem * pb->npe) pb->vbuf.alloc(pb->n_elem * pb->npe);
        return;
    }
    if (!use_gather(pb) || !uniform_geometry(pb)) return;
    const int64_t npe = pb->npe, m = pb->n_elem * npe;
    if (pb->uhat.n < 2 * npe * npe) pb->uhat.alloc(2 * npe * npe);
    if (pb->fbuf.n < m) pb->fbuf.alloc(m);
    if (pb->vbuf.n < m) pb->vbuf.alloc(m);
}

// what assembling these forms (with the mass vector) will allocate on the
// device, reserved now: callers that capture assembly into a graph call it
// first, since no allocation may happen during capture
void assembly_reserve(rsd_problem* pb, int32_t nf, const int32_t* kinds, bool with_mass) {
    AsmArgs A = base_args(pb);
    set_forms(A, nf, kinds, nullptr);
    A.need_state = forms_need_state(pb, A);
    if (use_uniform(pb, A)) {
        const int64_t npe = pb->npe;
        if (pb->uhat.n < 2 * npe * npe) pb->uhat.alloc(2 * npe * npe);
        if (with_mass && pb->fbuf.n < pb->n_elem * npe) pb->fbuf.alloc(pb->n_elem * npe);
    } else {
        gather_reserve(pb, with_mass);
    }
}

void assemble_matrix(rsd_problem* pb, double time, int32_t nf, const int32_t* kinds, const double* coefs,
                     const double* U, const double* Udot, rsd_matrix* K) {
    RSD_CHECK(K->pb == pb, "matrix belongs to another problem");
    AsmArgs A = base_args(pb);
    set_forms(A, nf, kinds, coefs);
    A.U = U;
    A.Udot = Udot;
    A.val = K->val.p;
    A.time = time;
    A.need_state = forms_need_state(pb, A);
    update_gval(pb, time);
    if (!use_gather(pb)) launch_fill(pb->ctx, K->val.n, K->val.p, 0.0);  // the row pass writes every slot
    dispatch_matrix(pb, A);
}

void assemble_diagonal(rsd_problem* pb, double time, int32_t nf, const int32_t* kinds, const double* coefs,
                       const double* U, double* diag) {
    AsmArgs A = base_args(pb);
    set_forms(A, nf, kinds, coefs);
    A.U = U;
    A.Udot = nullptr;  // the rate enters no diagonal term of the heat forms
    A.F = diag;
    A.time = time;
    A.need_state = forms_need_state(pb, A);
    update_gval(pb, time);
    launch_fill(pb->ctx, pb->n_free, diag, 0.0);
    dispatch_matrix(pb, A);
}

bool matrix_mass_fusable(const rsd_problem* pb, int32_t nf, const int32_t* kinds) {
    AsmArgs A{};
    set_forms(A, nf, kinds, nullptr);
    return !forms_need_state(pb, A) || !pb->bc_time_dependent;
}

void assemble_matrix_mass(rsd_problem* pb, double time, int32_t nf, const int32_t* kinds, const double* coefs,
                          const double* U, const double* Udot, rsd_matrix* K, const double* Um, double time_m,
                          double* Fm, const double* axx, double* axo) {
    RSD_CHECK(K->pb == pb, "matrix belongs to another problem");
    RSD_CHECK(matrix_mass_fusable(pb, nf, kinds), "matrix and mass vector need Dirichlet values at two times");
    RSD_CHECK(!axo || (axx && gather_forms_product(pb)), "the product with the matrix needs the row gather pass");
    AsmArgs A = base_args(pb);
    set_forms(A, nf, kinds, coefs);
    A.U = U;
    A.Udot = Udot;
    A.val = K->val.p;
    A.time = time;
    A.need_state = forms_need_state(pb, A);
    A.mvU = Um;
    A.mvF = Fm;
    A.axx = axx;
    A.axo = axo;
    // the state gather (if any) and the mass gather read the same values
    update_gval(pb, A.need_state ? time : time_m);
    if (!use_gather(pb)) {  // the row pass writes every slot and every entry of Fm
        launch_fill(pb->ctx, K->val.n, K->val.p, 0.0);
        launch_fill(pb->ctx, pb->n_free, Fm, 0.0);
    }
    dispatch_matrix(pb, A);
}

__global__ void k_zero_unless(int64_t n, double* __restrict__ p, const int32_t* __restrict__ stop) {
    if (*stop) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

void assemble_vector(rsd_problem* pb, double time, int32_t nf, const int32_t* kinds, const double* coefs,
                     const double* U, const double* Udot, const double* operand, int32_t selector, double* F,
                     const int32_t* stop) {
    RSD_CHECK(selector >= 0 && selector <= 2, "bad selector");
    AsmArgs A = base_args(pb);
    A.stop = stop;
    set_forms(A, nf, kinds, coefs);
    A.U = U;
    A.Udot = Udot;
    A.operand = operand;
    A.selector = selector;
    A.F = F;
    A.time = time;
    A.need_state = forms_need_state(pb, A);
    update_gval(pb, time);
    if (stop) {
        if (pb->n_free)
            hipLaunchKernelGGL(k_zero_unless, dim3((unsigned)nblocks(pb->n_free)), dim3(kBlock), 0, pb->ctx->stream,
                               pb->n_free, F, stop);
    } else {
        launch_fill(pb->ctx, pb->n_free, F, 0.0);
    }
    dispatch_vector(pb, A, Expr{}, 0);
}

void assemble_load(rsd_problem* pb, double time, const Expr* src, const double* nodal, double* F) {
    launch_fill(pb->ctx, pb->n_free, F, 0.0);
    if (!nodal && !src) return;  // a zero source adds exact zeros
    AsmArgs A = base_args(pb);
    A.F = F;
    A.nodal = nodal;
    A.time = time;
    dispatch_vector(pb, A, src ? *src : Expr{}, 1);
}

void apply_essential(rsd_problem* pb, double time, int32_t nf, const int32_t* kinds, const double* coefs, double* F) {
    AsmArgs A = base_args(pb);
    set_forms(A, nf, kinds, coefs);
    A.F = F;
    A.time = time;
    A.need_state = forms_need_state(pb, A);
    update_gval(pb, time);
    dispatch_vector(pb, A, Expr{}, 2);
}

void adjoint_grad_step(rsd_problem* pb, double time, const double* U, const double* lam, const double* kappa,
                       double* grad) {
    AsmArgs A = base_args(pb);
    A.U = U;
    A.operand = lam;
    A.time = time;
    update_gval(pb, time);
    // the element diffusion matrix with the per-element factor taken out:
    // kappa_e * Khat_e is the forward's element matrix, so Khat_e keeps the
    // model's k_scalar and tensor
    Model unit = pb->model;
    unit.has_k = 0;
    unit.has_dk = 0;
    if (pb->dim == 2 && pb->p == 1 && pb->nq == 2) { run_adj<2, 1, 2>(pb, A, unit, kappa, grad); return; }
    if (pb->dim == 2 && pb->p == 2 && pb->nq == 3) { run_adj<2, 2, 3>(pb, A, unit, kappa, grad); return; }
    if (pb->dim == 3 && pb->p == 1 && pb->nq == 2) { run_adj<3, 1, 2>(pb, A, unit, kappa, grad); return; }
    throw Error(RSD_ERR_ARG, "adjoint gradient: discretisation not supported (Q1/Q2 quad, Q1 hex, nq = p + 1)");
}

const Expr* cached_source(rsd_problem* pb, const char* text) {
    if (!text) return nullptr;
    if (pb->src_text != text) {
        pb->src_expr = compile_expr(text);
        pb->src_text = text;
    }
    // an expression that is the literal 0 contributes exact zeros
    if (pb->src_expr.n == 1 && pb->src_expr.op[0] == E_CONST && pb->src_expr.c[0] == 0.0) return nullptr;
    return &pb->src_expr;
}

}  // namespace rsd

using namespace rsd;

extern "C" {

int rsd_assemble_matrix(rsd_ctx* ctx, rsd_problem* pb, double time, int32_t nf, const int32_t* kinds,
                        const double* coefs, const double* U, const double* Udot, rsd_matrix* K) {
    return guard([&] {
        RSD_CHECK(ctx && pb && K && (nf == 0 || kinds), "rsd_assemble_matrix: null argument");
        RSD_LIVE_MATRIX(K);
        RSD_CHECK(ctx == pb->ctx, "problem belongs to another context");
        assemble_matrix(pb, time, nf, kinds, coefs, U, Udot, K);
    });
}

int rsd_assemble_diagonal(rsd_ctx* ctx, rsd_problem* pb, double time, int32_t nf, const int32_t* kinds,
                          const double* coefs, const double* U, double* diag) {
    return guard([&] {
        RSD_CHECK(ctx && pb && diag && (nf == 0 || kinds), "rsd_assemble_diagonal: null argument");
        RSD_CHECK(ctx == pb->ctx, "problem belongs to another context");
        assemble_diagonal(pb, time, nf, kinds, coefs, U, diag);
    });
}

int rsd_assemble_vector(rsd_ctx* ctx, rsd_problem* pb, double time, int32_t nf, const int32_t* kinds,
                        const double* coefs, const double* U, const double* Udot, const double* operand,
                        int32_t selector, double* F) {
    return guard([&] {
        RSD_CHECK(ctx && pb && F && (nf == 0 || kinds), "rsd_assemble_vector: null argument");
        RSD_CHECK(ctx == pb->ctx, "problem belongs to another context");
        assemble_vector(pb, time, nf, kinds, coefs, U, Udot, operand, selector, F);
    });
}

int rsd_assemble_load(rsd_ctx* ctx, rsd_problem* pb, double time, const char* src_expr, const double* nodal,
                      double* F) {
    return guard([&] {
        RSD_CHECK(ctx && pb && F, "rsd_assemble_load: null argument");
        RSD_CHECK(ctx == pb->ctx, "problem belongs to another context");
        assemble_load(pb, time, nodal ? nullptr : cached_source(pb, src_expr), nodal, F);
    });
}

int rsd_apply_essential(rsd_ctx* ctx, rsd_problem* pb, double time, int32_t nf, const int32_t* kinds,
                        const double* coefs, double* F) {
    return guard([&] {
        RSD_CHECK(ctx && pb && F && (nf == 0 || kinds), "rsd_apply_essential: null argument");
        RSD_CHECK(ctx == pb->ctx, "problem belongs to another context");
        apply_essential(pb, time, nf, kinds, coefs, F);
    });
}

}  // extern "C"
