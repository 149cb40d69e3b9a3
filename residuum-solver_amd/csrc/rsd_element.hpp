// rsd_element.hpp -- element geometry and heat-form integrands shared by
// the element kernels (assembly, adjoint gradient, boundary reaction).
//
// The element math follows the reference operand by operand
// (EvalQuadraturePointVolume.hpp:64-85, JacobianTransform.tpp:3-206,
// Lagrange1D.tpp:4-51, the heat-equation forms) with every product and sum
// rounded separately.
#pragma once

#include "rsd_assembly.hpp"

namespace rsd {

template<int P>
__device__ __forceinline__ void lag1d(double xi, double* N, double* dN) {
    if constexpr (P == 1) {
        N[0] = 0.5 * (1.0 - xi); N[1] = 0.5 * (1.0 + xi);
        dN[0] = -0.5; dN[1] = 0.5;
    } else {
        N[0] = 0.5 * xi * (xi - 1.0); N[1] = (1.0 - xi) * (1.0 + xi); N[2] = 0.5 * xi * (xi + 1.0);
        dN[0] = xi - 0.5; dN[1] = -2.0 * xi; dN[2] = xi + 0.5;
    }
}

// Gauss-Legendre nodes/weights (GaussQuadrature1D.tpp:4-71)
template<int NQ>
__device__ __forceinline__ void gauss(int i, double& x, double& w) {
    if constexpr (NQ == 1) { x = 0.0; w = 2.0; }
    else if constexpr (NQ == 2) { x = i == 0 ? -0.5773502691896257 : 0.5773502691896257; w = 1.0; }
    else if constexpr (NQ == 3) {
        x = i == 0 ? -0.7745966692414834 : (i == 1 ? 0.0 : 0.7745966692414834);
        w = i == 1 ? 0.8888888888888889 : 0.5555555555555556;
    } else {
        const double X[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
        const double W[4] = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538};
        x = X[i]; w = W[i];
    }
}

template<int D, int P>
struct Elem {
    static constexpr int NPE = D == 2 ? (P + 1) * (P + 1) : (P + 1) * (P + 1) * (P + 1);
};

// quadrature-point data (EvalQuadraturePointVolume members)
template<int D, int NPE>
struct QP {
    double N[NPE], dNdx[NPE * D], x[D], measure, w;
    double T, Trate, gradT[D], K[D * D], dKdT[D * D], rho, cp, dcpdT;
};

// Geometry at the reference point xi: N, x (need_x), J (into Jout when
// given), the measure sqrt(det g) and, GRAD, dN/dx = J^-1 dN/dxi.  GRAD =
// false leaves out J^-1 and dN/dx (forms without gradients); the values
// computed are the same either way.
template<int D, int P, bool GRAD = true>
__device__ __forceinline__ void qp_at(const double (&xi)[3], const double (&X)[Elem<D, P>::NPE][D],
                                      QP<D, Elem<D, P>::NPE>& Q, bool need_x, double* Jout = nullptr) {
    constexpr int NPE = Elem<D, P>::NPE;
    double Nx[D][P + 1], dNx[D][P + 1];
#pragma unroll
    for (int k = 0; k < D; ++k) lag1d<P>(xi[k], Nx[k], dNx[k]);
    double dNdxi[NPE * D];
#pragma unroll
    for (int a = 0; a < NPE; ++a) {
        const int i = a % (P + 1), j = (a / (P + 1)) % (P + 1), k = a / ((P + 1) * (P + 1));
        if constexpr (D == 2) {
            Q.N[a] = Nx[0][i] * Nx[1][j];
            dNdxi[a * 2] = dNx[0][i] * Nx[1][j];
            dNdxi[a * 2 + 1] = Nx[0][i] * dNx[1][j];
        } else {
            Q.N[a] = Nx[0][i] * Nx[1][j] * Nx[2][k];
            dNdxi[a * 3] = dNx[0][i] * Nx[1][j] * Nx[2][k];
            dNdxi[a * 3 + 1] = Nx[0][i] * dNx[1][j] * Nx[2][k];
            dNdxi[a * 3 + 2] = Nx[0][i] * Nx[1][j] * dNx[2][k];
        }
    }
    if (need_x) {
#pragma unroll
        for (int i = 0; i < D; ++i) Q.x[i] = 0.0;
#pragma unroll
        for (int b = 0; b < NPE; ++b)
#pragma unroll
            for (int i = 0; i < D; ++i) Q.x[i] += Q.N[b] * X[b][i];
    }
    double J[D * D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int al = 0; al < D; ++al) {
            double s = 0.0;
#pragma unroll
            for (int b = 0; b < NPE; ++b) s += dNdxi[b * D + al] * X[b][i];
            J[i * D + al] = s;
        }
    if (Jout) {
#pragma unroll
        for (int i = 0; i < D * D; ++i) Jout[i] = J[i];
    }
    double g[D * D];
#pragma unroll
    for (int be = 0; be < D; ++be)
#pragma unroll
        for (int al = 0; al < D; ++al) {
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < D; ++i) s += J[i * D + al] * J[i * D + be];
            g[be * D + al] = s;
            g[al * D + be] = s;
        }
    double inv[D * D];
    if constexpr (D == 2) {
        Q.measure = sqrt((g[0] * g[3]) - (g[1] * g[2]));
        if constexpr (!GRAD) return;
        const double di = 1.0 / ((J[0] * J[3]) - (J[1] * J[2]));
        inv[0] = di * J[3]; inv[1] = -di * J[1]; inv[2] = -di * J[2]; inv[3] = di * J[0];
    } else {
        const double dg = (g[0] * ((g[4] * g[8]) - (g[5] * g[7]))) - (g[1] * ((g[3] * g[8]) - (g[5] * g[6]))) +
                          (g[2] * ((g[3] * g[7]) - (g[4] * g[6])));
        Q.measure = sqrt(dg);
        if constexpr (!GRAD) return;
        const double* A = J;
        const double dj = (A[0] * ((A[4] * A[8]) - (A[5] * A[7]))) - (A[1] * ((A[3] * A[8]) - (A[5] * A[6]))) +
                          (A[2] * ((A[3] * A[7]) - (A[4] * A[6])));
        const double di = 1.0 / dj;
        inv[0] = di * ((A[4] * A[8]) - (A[5] * A[7]));
        inv[1] = -di * ((A[3] * A[8]) - (A[5] * A[6]));
        inv[2] = di * ((A[3] * A[7]) - (A[4] * A[6]));
        inv[3] = -di * ((A[1] * A[8]) - (A[2] * A[7]));
        inv[4] = di * ((A[0] * A[8]) - (A[2] * A[6]));
        inv[5] = -di * ((A[0] * A[7]) - (A[1] * A[6]));
        inv[6] = di * ((A[1] * A[5]) - (A[2] * A[4]));
        inv[7] = -di * ((A[0] * A[5]) - (A[2] * A[3]));
        inv[8] = di * ((A[0] * A[4]) - (A[1] * A[3]));
    }
#pragma unroll
    for (int b = 0; b < NPE; ++b)
#pragma unroll
        for (int i = 0; i < D; ++i) {
            double s = 0.0;
#pragma unroll
            for (int al = 0; al < D; ++al) s += inv[i * D + al] * dNdxi[b * D + al];
            Q.dNdx[b * D + i] = s;
        }
}

// volume point q of the tensor Gauss rule (x fastest)
template<int D, int P, int NQ, bool GRAD = true>
__device__ __forceinline__ void qp_eval(int q, const double (&X)[Elem<D, P>::NPE][D], QP<D, Elem<D, P>::NPE>& Q,
                                        bool need_x) {
    const int qi = q % NQ, qj = (q / NQ) % NQ, qk = q / (NQ * NQ);
    double xi[3] = {0.0, 0.0, 0.0}, wq[3];
    gauss<NQ>(qi, xi[0], wq[0]);
    gauss<NQ>(qj, xi[1], wq[1]);
    if constexpr (D == 3) gauss<NQ>(qk, xi[2], wq[2]);
    Q.w = D == 2 ? wq[0] * wq[1] : wq[0] * wq[1] * wq[2];
    qp_at<D, P, GRAD>(xi, X, Q, need_x);
}

template<int D, int NPE>
__device__ __forceinline__ void qp_state(QP<D, NPE>& Q, const double* Uv, const double* Ur) {
    double u = 0.0;
#pragma unroll
    for (int a = 0; a < NPE; ++a) u += Q.N[a] * Uv[a];
    Q.T = u;
#pragma unroll
    for (int i = 0; i < D; ++i) Q.gradT[i] = 0.0;
#pragma unroll
    for (int a = 0; a < NPE; ++a)
#pragma unroll
        for (int i = 0; i < D; ++i) Q.gradT[i] += Q.dNdx[a * D + i] * Uv[a];
    double r = 0.0;
#pragma unroll
    for (int a = 0; a < NPE; ++a) r += Q.N[a] * Ur[a];
    Q.Trate = r;
}

template<int D, int P>
__device__ __forceinline__ void load_element(const AsmArgs& A, int64_t e, int32_t (&nodes)[Elem<D, P>::NPE],
                                             double (&X)[Elem<D, P>::NPE][D]) {
    constexpr int NPE = Elem<D, P>::NPE;
#pragma unroll
    for (int a = 0; a < NPE; ++a) {
        nodes[a] = A.ien[e * NPE + a];
#pragma unroll
        for (int i = 0; i < D; ++i) X[a][i] = A.xyz[(int64_t)nodes[a] * D + i];
    }
}

// gather modes: 0 free, 1 full (free values + Dirichlet values), 2 constrained
template<int NPE>
__device__ __forceinline__ void gather(const AsmArgs& A, const int32_t (&nodes)[NPE], int mode, const double* U,
                                       double (&Ue)[NPE]) {
#pragma unroll
    for (int a = 0; a < NPE; ++a) {
        const int32_t r = A.t2a[nodes[a]];
        double v = 0.0;
        if (r >= 0) { if (mode != 2 && U) v = U[r]; }
        else if (mode != 0) v = A.gval[nodes[a]];
        Ue[a] = v;
    }
}

// ------------------------------------------------------------------------
// Element kernels: NPE threads per element.
//   phase 1  the element's threads split its quadrature points; each
//            evaluates the geometry (and, for state-dependent models, the
//            state and the conductivity) of its points into LDS;
//   phase 2  thread a accumulates row a of the element matrix (entry a of
//            the element vector) over points, forms and columns in the
//            reference's order, then scatters it.
// Each entry's accumulation chain is the one a single thread per element
// would run, so results are bitwise those of the serial loops.
// ------------------------------------------------------------------------
template<int D, int P, int NQ, bool STATE, bool GRAD = true>
struct ECfg {
    static_assert(GRAD || !STATE, "state-dependent models need the gradients");
    static constexpr int NPE = Elem<D, P>::NPE;
    static constexpr int NQP = D == 2 ? NQ * NQ : NQ * NQ * NQ;
    // per-point record: dNdx[NPE*D] (GRAD), measure, source value, then
    // (state) k, dk, Trate, gradT[D]
    static constexpr int R_DN = 0, R_MEAS = GRAD ? NPE * D : 0, R_SRC = R_MEAS + 1;
    static constexpr int R_K = R_SRC + 1, R_DK = R_K + 1, R_TR = R_DK + 1, R_GT = R_TR + 1;
    static constexpr int REC = STATE ? R_GT + D : R_SRC + 1;
    // element records ELEM_DOUBLES apart, a stride of 2 mod 16 doubles (4
    // mod 32 dwords): the 8 elements of a wavefront, whose threads read the
    // same record offset in one instruction, start on 8 distinct groups of 4
    // LDS banks instead of one (an unpadded stride that is a multiple of 16
    // doubles puts them all on the same bank)
    static constexpr int ELEM_RAW = NQP * REC;
    static constexpr int ELEM_DOUBLES = ELEM_RAW + (18 - ELEM_RAW % 16) % 16;
    static constexpr int LDS_BUDGET = 40 * 1024;
    static constexpr int EPB_RAW = LDS_BUDGET / (ELEM_DOUBLES * 8);
    static constexpr int EPB_MAX = 256 / NPE;
    static constexpr int EPB_CAP = EPB_RAW < EPB_MAX ? EPB_RAW : EPB_MAX;
    // keep whole wavefronts when the element size allows it
    static constexpr int WAVE_EL = 64 % NPE == 0 ? 64 / NPE : 1;
    static constexpr int EPB_W = EPB_CAP >= WAVE_EL ? (EPB_CAP / WAVE_EL) * WAVE_EL : EPB_CAP;
    static constexpr int EPB = EPB_W > 0 ? EPB_W : 1;
    static constexpr int BLOCK = EPB * NPE;
};

template<int D, int P>
__device__ __forceinline__ void shape_values(const double (&xi)[3], double (&N)[Elem<D, P>::NPE]) {
    constexpr int NPE = Elem<D, P>::NPE;
    double Nx[D][P + 1], dNx[D][P + 1];
#pragma unroll
    for (int k = 0; k < D; ++k) lag1d<P>(xi[k], Nx[k], dNx[k]);
#pragma unroll
    for (int a = 0; a < NPE; ++a) {
        const int i = a % (P + 1), j = (a / (P + 1)) % (P + 1), k = a / ((P + 1) * (P + 1));
        if constexpr (D == 2) N[a] = Nx[0][i] * Nx[1][j];
        else N[a] = Nx[0][i] * Nx[1][j] * Nx[2][k];
    }
}

template<int D, int NQ>
__device__ __forceinline__ double point_coords(int q, double (&xi)[3]) {
    const int qi = q % NQ, qj = (q / NQ) % NQ, qk = q / (NQ * NQ);
    double wq[3];
    gauss<NQ>(qi, xi[0], wq[0]);
    gauss<NQ>(qj, xi[1], wq[1]);
    xi[2] = 0.0;
    if constexpr (D == 3) {
        gauss<NQ>(qk, xi[2], wq[2]);
        return wq[0] * wq[1] * wq[2];
    } else {
        return wq[0] * wq[1];
    }
}

// phase 1: geometry (and state) of the points q = a, a + NPE, ... of one element
//   load: 0 none, 1 source expression at x_q, 2 nodal source interpolated
template<int D, int P, int NQ, bool STATE, bool GRAD = true>
__device__ __forceinline__ void phase1(const AsmArgs& A, const Model& m, const Expr* src, int load, int64_t e, int a,
                                       const double (&X)[Elem<D, P>::NPE][D], const double* Uv, const double* Ur,
                                       const double* Un, double kap, bool has_kap, double* rec) {
    using C = ECfg<D, P, NQ, STATE, GRAD>;
    constexpr int NPE = C::NPE;
    for (int q = a; q < C::NQP; q += NPE) {
        QP<D, NPE> Q;
        qp_eval<D, P, NQ, GRAD>(q, X, Q, load == 1);
        double* r = rec + q * C::REC;
        if constexpr (GRAD) {
#pragma unroll
            for (int i = 0; i < NPE * D; ++i) r[C::R_DN + i] = Q.dNdx[i];
        }
        r[C::R_MEAS] = Q.measure;
        if (load == 1) {
            r[C::R_SRC] = eval_expr(*src, A.time, Q.x[0], Q.x[1], D == 3 ? Q.x[2] : 0.0, 0.0);
        } else if (load == 2) {
            double val = 0.0;
#pragma unroll
            for (int b = 0; b < NPE; ++b) val += Q.N[b] * Un[b];
            r[C::R_SRC] = val;
        }
        if constexpr (STATE) {
            double Uvl[NPE], Url[NPE];
#pragma unroll
            for (int b = 0; b < NPE; ++b) { Uvl[b] = Uv[b]; Url[b] = Ur[b]; }
            qp_state<D, NPE>(Q, Uvl, Url);
            double k = m.has_k ? eval_expr(m.k, 0.0, 0.0, 0.0, 0.0, Q.T) : m.k_scalar;
            double dk = m.has_dk ? eval_expr(m.dk, 0.0, 0.0, 0.0, 0.0, Q.T) : 0.0;
            if (has_kap) { k = kap * k; dk = kap * dk; }
            r[C::R_K] = k;
            r[C::R_DK] = dk;
            r[C::R_TR] = Q.Trate;
#pragma unroll
            for (int i = 0; i < D; ++i) r[C::R_GT + i] = Q.gradT[i];
        }
    }
}

// conductivity tensor entry (qp_model: K = k * shape)
__device__ __forceinline__ double kshape(const Model& m, int D, int i, int j) {
    return m.aniso ? m.tensor[i * D + j] : ((i == j) ? 1.0 : 0.0);
}

// diffusion integrand dNdx_a . (K dNdx_b), K = k * shape; for an isotropic
// model the zero off-diagonal products are left out (they only add zeros)
template<int D>
__device__ __forceinline__ double diff_integrand(const Model& m, double k, const double* da, const double* db) {
    double integrand = 0.0;
    if (!m.aniso) {
#pragma unroll
        for (int i = 0; i < D; ++i) integrand += da[i] * (k * db[i]);
    } else {
#pragma unroll
        for (int i = 0; i < D; ++i) {
            double mv = 0.0;
#pragma unroll
            for (int j = 0; j < D; ++j) mv += (k * m.tensor[i * D + j]) * db[j];
            integrand += da[i] * mv;
        }
    }
    return integrand;
}

// dNdx_a . (dKdT gradT), dKdT = dk * shape
template<int D>
__device__ __forceinline__ double tdiff_integrand(const Model& m, double dk, const double* da, const double* gt) {
    double integrand = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        double mv = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) mv += (dk * kshape(m, D, i, j)) * gt[j];
        integrand += da[i] * mv;
    }
    return integrand;
}

}  // namespace rsd
