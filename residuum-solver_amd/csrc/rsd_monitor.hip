// rsd_monitor.hip -- boundary heat-flux monitors per boundary tag.
//
// Weak reaction: QuantityEvaluator<CPU>::evaluateBoundaryReaction
// (fem/quantity/backend/cpu/QuantityEvaluator.tpp:95-199) with the heat
// problem's pairs (HeatProblem.tpp:385-410): every element with a face that
// carries a requested tag forms Oe = source - K(U) U - M dU/dt, the three
// contributions interleaved per quadrature point; each tagged face then adds
// the sum of Oe over its nodes, a node shared by two faces of the same tag
// counted once per element.
// Strong flux: QuantityEvaluator<CPU>::evaluateBoundaryRegistry
// (QuantityEvaluator.tpp:14-93) with HeatFluxIntegrand: each tagged face
// adds -(K gradT).n w over its points.
// Both add the faces' measure, the sum over face points of |cof(J) n_ref| w.
//
// Device work:
//   k_bdy_element  Oe of every boundary element (NPE threads per element,
//                  the volume points split over them as in the assembly
//                  kernels; outputs are private to the element, so no
//                  colouring);
//   k_bdy_faces    one thread per face point: the measure term and, strong,
//                  the flux term;
//   k_ordered_sums each tag's terms summed in the serial loop's order
//                  (element, face, face node / face point) by one wavefront
//                  per sum: lanes load 64 terms at a time and the running sum
//                  takes them in lane order, so every sum is bitwise the
//                  serial one.
// The plan (boundary elements, the claimed Oe entries and the face slots of
// every tag, in loop order) depends only on the mesh and the tag list and is
// cached on the problem.
#include "rsd_assembly.hpp"
#include "rsd_element.hpp"
#include "rsd_face.hpp"

#include <algorithm>

namespace rsd {

struct ReactionPlan {
    std::vector<int32_t> tags;
    int64_t nbe = 0, n_points = 0;
    std::vector<int32_t> h_belem, h_sface, h_idx;
    std::vector<int64_t> h_roff, h_foff;
    DBuf<int32_t> belem;  // [nbe] boundary elements, ascending
    DBuf<int32_t> sface;  // [n_slots] tagged face of each slot, boundary element * nfaces + face, grouped by tag
    DBuf<int32_t> idx;    // claimed entries of oe, grouped by tag in loop order
    DBuf<int64_t> roff;   // [n_tags + 1] per-tag ranges of idx
    DBuf<int64_t> foff;   // [n_tags + 1] per-tag ranges of the face-point terms
    DBuf<double> oe;      // [nbe * npe]
    DBuf<double> mterm;   // [n_points] measure terms
    DBuf<double> fterm;   // [n_points] strong flux terms
    DBuf<double> out;     // [2 * n_tags] results, then measures
};

static ReactionPlan& reaction_plan(rsd_problem* pb, int32_t n_tags, const int32_t* tags) {
    std::vector<int32_t> key(tags, tags + n_tags);
    if (pb->reaction && pb->reaction->tags == key) return *pb->reaction;
    auto plan = std::make_shared<ReactionPlan>();
    ReactionPlan& R = *plan;
    R.tags = key;
    const int NF = pb->nfaces, npe = pb->npe, D = pb->dim;
    const int FQ = D == 2 ? pb->nq : pb->nq * pb->nq;
    std::vector<std::vector<int32_t>> fn(NF);
    for (int f = 0; f < NF; ++f) fn[f] = face_local_nodes(D, pb->p, f);
    std::vector<std::vector<int32_t>> ridx(n_tags);
    std::vector<std::vector<int64_t>> faces(n_tags);  // (boundary element, face) pairs per tag
    std::vector<int32_t> claimed(npe);
    for (int64_t e = 0; e < pb->n_elem; ++e) {
        const int32_t* rt = pb->h_rng.data() + e * NF;
        bool touches = false;
        for (int f = 0; f < NF && !touches; ++f)
            for (int t = 0; t < n_tags; ++t) touches |= rt[f] == tags[t];
        if (!touches) continue;
        const int64_t be = (int64_t)R.h_belem.size();
        R.h_belem.push_back((int32_t)e);
        std::fill(claimed.begin(), claimed.end(), -1);
        for (int f = 0; f < NF; ++f) {
            int t = -1;
            for (int k = 0; k < n_tags; ++k)
                if (rt[f] == tags[k]) t = k;
            if (t < 0) continue;
            faces[t].push_back(be * NF + f);
            for (int32_t a : fn[f])
                if (claimed[a] != rt[f]) {
                    ridx[t].push_back((int32_t)(be * npe + a));
                    claimed[a] = rt[f];
                }
        }
    }
    R.nbe = (int64_t)R.h_belem.size();
    RSD_CHECK(R.nbe * std::max(npe, NF) < INT32_MAX, "boundary monitors: too many boundary elements");
    R.h_roff.assign((size_t)n_tags + 1, 0);
    R.h_foff.assign((size_t)n_tags + 1, 0);
    for (int t = 0; t < n_tags; ++t) {
        R.h_roff[t + 1] = R.h_roff[t] + (int64_t)ridx[t].size();
        R.h_idx.insert(R.h_idx.end(), ridx[t].begin(), ridx[t].end());
        for (int64_t bf : faces[t]) R.h_sface.push_back((int32_t)bf);
        R.h_foff[t + 1] = (int64_t)R.h_sface.size() * FQ;
    }
    R.n_points = (int64_t)R.h_sface.size() * FQ;
    hipStream_t st = pb->ctx->stream;
    R.belem.upload(R.h_belem.data(), R.nbe, st);
    R.sface.upload(R.h_sface.data(), (int64_t)R.h_sface.size(), st);
    R.idx.upload(R.h_idx.data(), (int64_t)R.h_idx.size(), st);
    R.roff.upload(R.h_roff.data(), (int64_t)R.h_roff.size(), st);
    R.foff.upload(R.h_foff.data(), (int64_t)R.h_foff.size(), st);
    R.oe.alloc(R.nbe * npe);
    R.mterm.alloc(R.n_points);
    R.fterm.alloc(R.n_points);
    R.out.alloc(2 * (int64_t)n_tags);
    pb->reaction = plan;
    return R;
}

// One thread per face point of the plan's tagged faces (slot s, point q):
// the measure term |n| w into mterm[s*FQ + q] and, FLUX, the strong flux
// term -(K gradT).n w (HeatFluxIntegrand.hpp:19-38) with T the full gather
// of A.U and K = k(T) shape into fterm[s*FQ + q].
template<int D, int P, int NQ, bool FLUX, bool STATE>
__global__ __launch_bounds__(kBlock) void k_bdy_faces(AsmArgs A, Model m, const int32_t* __restrict__ belem,
                                                      const int32_t* __restrict__ sface, int64_t n_points,
                                                      double* __restrict__ mterm, double* __restrict__ fterm) {
    constexpr int NPE = Elem<D, P>::NPE, NF = 2 * D, FQ = D == 2 ? NQ : NQ * NQ;
    const int64_t tid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (tid >= n_points) return;
    const int64_t s = tid / FQ;
    const int q = (int)(tid % FQ);
    const int32_t bf = sface[s];
    const int f = bf % NF;
    const int64_t e = belem[bf / NF];
    int32_t nodes[NPE];
    double X[NPE][D];
    load_element<D, P>(A, e, nodes, X);
    double xi[3], nref[D];
    const double fw = face_point<D, NQ>(f, q, xi, nref);
    QP<D, NPE> Q;
    double J[D * D];
    qp_at<D, P, FLUX>(xi, X, Q, false, J);
    double n[D];
    boundary_normal<D>(J, nref, n);
    double mag = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) mag += n[i] * n[i];
    mterm[tid] = sqrt(mag) * fw;
    if constexpr (FLUX) {
        double Uv[NPE], Ur[NPE];
        gather<NPE>(A, nodes, 1, A.U, Uv);
#pragma unroll
        for (int a = 0; a < NPE; ++a) Ur[a] = 0.0;
        qp_state<D, NPE>(Q, Uv, Ur);
        double k = STATE && m.has_k ? eval_expr(m.k, 0.0, 0.0, 0.0, 0.0, Q.T) : m.k_scalar;
        if (A.kappa) k = A.kappa[e] * k;
        double flux = 0.0;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            double kg = 0.0;
#pragma unroll
            for (int j = 0; j < D; ++j) kg += (k * kshape(m, D, i, j)) * Q.gradT[j];
            flux -= kg * n[i];
        }
        fterm[tid] = flux * fw;
    }
}

template<int D, int P, int NQ, bool STATE>
__global__ __launch_bounds__((ECfg<D, P, NQ, STATE>::BLOCK)) void k_bdy_element(
    AsmArgs A, Model m, Expr src, int32_t has_src, const int32_t* __restrict__ belem, int64_t nbe,
    double* __restrict__ oe) {
    using C = ECfg<D, P, NQ, STATE>;
    constexpr int NPE = C::NPE;
    __shared__ double lds[C::EPB * C::ELEM_DOUBLES];
    const int el = threadIdx.x / NPE, a = threadIdx.x % NPE;
    const int64_t k = (int64_t)blockIdx.x * C::EPB + el;
    const bool valid = k < nbe;
    double* rec = lds + el * C::ELEM_DOUBLES;
    double kap = 1.0;
    const bool has_kap = A.kappa != nullptr;
    double Uv[NPE], Ur[NPE];
    if (valid) {
        const int64_t e = belem[k];
        int32_t nodes[NPE];
        double X[NPE][D];
        load_element<D, P>(A, e, nodes, X);
        if (has_kap) kap = A.kappa[e];
        gather<NPE>(A, nodes, 1, A.U, Uv);     // Full: free values and Dirichlet values
        gather<NPE>(A, nodes, 0, A.Udot, Ur);  // Free
        phase1<D, P, NQ, STATE>(A, m, &src, has_src ? 1 : 0, e, a, X, Uv, Ur, nullptr, kap, has_kap, rec);
    }
    __syncthreads();
    if (!valid) return;
    double kel = m.k_scalar;
    if (has_kap) kel = kap * kel;
    double Oe = 0.0;
    for (int q = 0; q < C::NQP; ++q) {
        const double* r = rec + q * C::REC;
        double xi[3];
        const double w = point_coords<D, NQ>(q, xi);
        double N[NPE];
        shape_values<D, P>(xi, N);
        const double meas = r[C::R_MEAS];
        // SourceForm on the value (SourceForm.hpp:22-32)
        if (has_src) Oe += r[C::R_SRC] * N[a] * meas * w;
        // ScaledForm(-1) of the stiffness forms on the value
        const double* da = r + C::R_DN + a * D;
        const double kq = STATE ? r[C::R_K] : kel;
        double buf = 0.0;
#pragma unroll
        for (int b = 0; b < NPE; ++b) buf += diff_integrand<D>(m, kq, da, r + C::R_DN + b * D) * Uv[b] * meas * w;
        Oe += -1.0 * buf;
        // ScaledForm(-1) of the mass forms on the rate (transient problems)
        if (A.Udot) {
            const double ra = m.rho * m.cp * N[a];
            double mb = 0.0;
#pragma unroll
            for (int b = 0; b < NPE; ++b) mb += ra * N[b] * Ur[b] * meas * w;
            Oe += -1.0 * mb;
        }
    }
    oe[k * NPE + a] = Oe;
}

__device__ __forceinline__ double lane_value(double v, int k) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), k);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), k);
    return __hiloint2double(hi, lo);
}

// one wavefront per sum g: out[g] = the terms vals[idx[j]] (idx given) or
// vals[j], j = off[g] .. off[g+1] - 1, added one after another from 0.0
__global__ __launch_bounds__(64) void k_ordered_sums(const double* __restrict__ vals, const int32_t* __restrict__ idx,
                                                     const int64_t* __restrict__ off, double* __restrict__ out) {
    const int g = blockIdx.x, lane = threadIdx.x;
    const int64_t j0 = off[g], j1 = off[g + 1];
    double acc = 0.0;
    for (int64_t base = j0; base < j1; base += 64) {
        const int64_t j = base + lane;
        double v = 0.0;
        if (j < j1) v = idx ? vals[idx[j]] : vals[j];
        const int cnt = (int)(j1 - base < 64 ? j1 - base : 64);
        for (int kk = 0; kk < cnt; ++kk) acc += lane_value(v, kk);
    }
    if (lane == 0) out[g] = acc;
}

enum MonitorKind { MON_WEAK = 0, MON_STRONG = 1 };

template<int D, int P, int NQ>
static void run_monitor(rsd_problem* pb, const AsmArgs& A, const Expr* src, ReactionPlan& R, int kind) {
    const bool state = pb->model.has_k || pb->model.has_dk;
    hipStream_t st = pb->ctx->stream;
    if (kind == MON_WEAK && R.nbe) {
        const Expr e = src ? *src : Expr{};
        const int32_t has_src = src != nullptr;
        if (state) {
            using C = ECfg<D, P, NQ, true>;
            hipLaunchKernelGGL((k_bdy_element<D, P, NQ, true>), dim3((unsigned)((R.nbe + C::EPB - 1) / C::EPB)),
                               dim3(C::BLOCK), 0, st, A, pb->model, e, has_src, R.belem.p, R.nbe, R.oe.p);
        } else {
            using C = ECfg<D, P, NQ, false>;
            hipLaunchKernelGGL((k_bdy_element<D, P, NQ, false>), dim3((unsigned)((R.nbe + C::EPB - 1) / C::EPB)),
                               dim3(C::BLOCK), 0, st, A, pb->model, e, has_src, R.belem.p, R.nbe, R.oe.p);
        }
    }
    if (R.n_points) {
        const dim3 grid((unsigned)nblocks(R.n_points));
        if (kind == MON_WEAK)
            hipLaunchKernelGGL((k_bdy_faces<D, P, NQ, false, false>), grid, dim3(kBlock), 0, st, A, pb->model,
                               R.belem.p, R.sface.p, R.n_points, R.mterm.p, R.fterm.p);
        else if (state)
            hipLaunchKernelGGL((k_bdy_faces<D, P, NQ, true, true>), grid, dim3(kBlock), 0, st, A, pb->model,
                               R.belem.p, R.sface.p, R.n_points, R.mterm.p, R.fterm.p);
        else
            hipLaunchKernelGGL((k_bdy_faces<D, P, NQ, true, false>), grid, dim3(kBlock), 0, st, A, pb->model,
                               R.belem.p, R.sface.p, R.n_points, R.mterm.p, R.fterm.p);
    }
    RSD_HIP(hipGetLastError());
}

static void dispatch_monitor(rsd_problem* pb, const AsmArgs& A, const Expr* src, ReactionPlan& R, int kind) {
    if (pb->dim == 2 && pb->p == 1 && pb->nq == 1) { run_monitor<2, 1, 1>(pb, A, src, R, kind); return; }
    if (pb->dim == 2 && pb->p == 1 && pb->nq == 2) { run_monitor<2, 1, 2>(pb, A, src, R, kind); return; }
    if (pb->dim == 2 && pb->p == 1 && pb->nq == 3) { run_monitor<2, 1, 3>(pb, A, src, R, kind); return; }
    if (pb->dim == 2 && pb->p == 2 && pb->nq == 2) { run_monitor<2, 2, 2>(pb, A, src, R, kind); return; }
    if (pb->dim == 2 && pb->p == 2 && pb->nq == 3) { run_monitor<2, 2, 3>(pb, A, src, R, kind); return; }
    if (pb->dim == 2 && pb->p == 2 && pb->nq == 4) { run_monitor<2, 2, 4>(pb, A, src, R, kind); return; }
    if (pb->dim == 3 && pb->p == 1 && pb->nq == 1) { run_monitor<3, 1, 1>(pb, A, src, R, kind); return; }
    if (pb->dim == 3 && pb->p == 1 && pb->nq == 2) { run_monitor<3, 1, 2>(pb, A, src, R, kind); return; }
    if (pb->dim == 3 && pb->p == 1 && pb->nq == 3) { run_monitor<3, 1, 3>(pb, A, src, R, kind); return; }
    throw Error(RSD_ERR_ARG, "discretisation not supported by the HIP element kernels (dim/p/nq)");
}

// the monitor's terms on the device, the ordered sums, the results to the host
static void monitor(rsd_ctx* ctx, rsd_problem* pb, int kind, double time, const Expr* src, const double* U,
                    const double* Udot, int32_t n_tags, const int32_t* tags, double* result, double* measure) {
    for (int32_t t = 0; t < n_tags; ++t) RSD_CHECK(tags[t] >= 0, "boundary tags are >= 0");
    ReactionPlan& R = reaction_plan(pb, n_tags, tags);
    AsmArgs A = base_args(pb);
    A.U = U;
    A.Udot = Udot;
    A.time = time;
    update_gval(pb, time);
    dispatch_monitor(pb, A, src, R, kind);
    if (kind == MON_WEAK)
        hipLaunchKernelGGL(k_ordered_sums, dim3((unsigned)n_tags), dim3(64), 0, ctx->stream, R.oe.p, R.idx.p,
                           R.roff.p, R.out.p);
    else
        hipLaunchKernelGGL(k_ordered_sums, dim3((unsigned)n_tags), dim3(64), 0, ctx->stream, R.fterm.p, nullptr,
                           R.foff.p, R.out.p);
    hipLaunchKernelGGL(k_ordered_sums, dim3((unsigned)n_tags), dim3(64), 0, ctx->stream, R.mterm.p, nullptr, R.foff.p,
                       R.out.p + n_tags);
    RSD_HIP(hipGetLastError());
    std::vector<double> h(2 * (size_t)n_tags);
    RSD_HIP(hipMemcpyAsync(h.data(), R.out.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    RSD_HIP(hipStreamSynchronize(ctx->stream));
    for (int32_t t = 0; t < n_tags; ++t) {
        result[t] = h[t];
        measure[t] = h[n_tags + t];
    }
}

}  // namespace rsd

using namespace rsd;

extern "C" {

int rsd_boundary_reaction(rsd_ctx* ctx, rsd_problem* pb, double time, const char* src_expr, const double* U,
                          const double* Udot, int32_t n_tags, const int32_t* tags, double* result, double* measure) {
    return guard([&] {
        RSD_CHECK(ctx && pb && U && (n_tags == 0 || (tags && result && measure)),
                  "rsd_boundary_reaction: null argument");
        RSD_CHECK(ctx == pb->ctx, "problem belongs to another context");
        RSD_CHECK(n_tags >= 0, "n_tags must be >= 0");
        if (n_tags == 0) return;
        monitor(ctx, pb, MON_WEAK, time, cached_source(pb, src_expr), U, Udot, n_tags, tags, result, measure);
    });
}

int rsd_boundary_flux(rsd_ctx* ctx, rsd_problem* pb, double time, const double* U, int32_t n_tags,
                      const int32_t* tags, double* result, double* measure) {
    return guard([&] {
        RSD_CHECK(ctx && pb && U && (n_tags == 0 || (tags && result && measure)), "rsd_boundary_flux: null argument");
        RSD_CHECK(ctx == pb->ctx, "problem belongs to another context");
        RSD_CHECK(n_tags >= 0, "n_tags must be >= 0");
        if (n_tags == 0) return;
        monitor(ctx, pb, MON_STRONG, time, nullptr, U, nullptr, n_tags, tags, result, measure);
    });
}

}  // extern "C"
