// rsd_natural.hip -- natural (flux) boundary conditions on the device:
// BoundaryApplicator<CPU>::applyNaturalBCs (fem/boundary/backend/cpu/
// BoundaryApplicator.tpp:66-140) with the heat equation's flux forms,
// FluxBoundaryForm (equation/heateq/form/FluxBoundaryForm.hpp:11-38, a flux
// vector q(t, x) given by expressions) and NodalFluxForm
// (NodalFluxForm.hpp:12-55, nodal flux vectors interpolated with the face
// shape values).
//
// The reference walks the elements in index order and, per element, its
// faces in index order; a face whose tag has entries gets a face vector
//   Fe[a] = sum over face points q, the tag's entries in order, sD:
//           (q_sD N_a) n_sD w      (n the unnormalised cofactor normal)
// which is added into F at the face's free nodes.  Device work:
//   k_natural_faces  one thread per tagged face forms Fe in exactly that
//                    order into a face buffer;
//   k_rows_add       one thread per free row touched by a tagged face adds
//                    its face terms in face order to F[row],
// so F is bitwise the serial loop's.  The face plan (tagged faces in loop
// order, the rows' term lists) depends only on the mesh and the tag list and
// is cached on the problem.
#include "rsd_assembly.hpp"
#include "rsd_element.hpp"
#include "rsd_face.hpp"

#include <algorithm>
#include <map>

namespace rsd {

constexpr int kNatMaxEntries = 8;

struct NaturalPlan {
    std::vector<int32_t> tags;    // distinct tags, the key
    int64_t n_faces = 0, n_rows = 0;
    DBuf<int32_t> face;           // [n_faces] element * nfaces + face, loop order
    DBuf<int32_t> ftag;           // [n_faces] the face's tag
    DBuf<int32_t> row;            // [n_rows] free rows touched
    DBuf<int32_t> off;            // [n_rows + 1]
    DBuf<int32_t> idx;            // face slot * npf + a, per row in face order
    DBuf<double> fe;              // [n_faces * npf]
    DBuf<Expr> ex;                // entry expressions, refreshed per call
};

// entries: tag, and either the device nodal flux [n_nodes * dim] or the
// dim expressions at ex[k * dim ..]
struct NatEntries {
    int32_t n;
    int32_t tag[kNatMaxEntries];
    const double* nodal[kNatMaxEntries];
    const Expr* ex;
};

static NaturalPlan& natural_plan(rsd_problem* pb, const std::vector<int32_t>& tags) {
    auto& slot = pb->natural;
    if (slot && slot->tags == tags) return *slot;
    auto plan = std::make_shared<NaturalPlan>();
    NaturalPlan& R = *plan;
    R.tags = tags;
    const int NF = pb->nfaces, npe = pb->npe, D = pb->dim;
    RSD_CHECK(pb->n_elem * NF < (int64_t)INT32_MAX, "natural boundary conditions: face ids exceed 32 bits");
    std::vector<std::vector<int32_t>> fn(NF);
    for (int f = 0; f < NF; ++f) fn[f] = face_local_nodes(D, pb->p, f);
    const int npf = (int)fn[0].size();
    std::vector<int32_t> faces, ftags;
    std::map<int32_t, std::vector<int32_t>> rows;  // free row -> its terms in face order
    for (int64_t e = 0; e < pb->n_elem; ++e) {
        const int32_t* rt = pb->h_rng.data() + e * NF;
        for (int f = 0; f < NF; ++f) {
            if (rt[f] < 0 || std::find(tags.begin(), tags.end(), rt[f]) == tags.end()) continue;
            const int32_t s = (int32_t)faces.size();
            faces.push_back((int32_t)(e * NF + f));
            ftags.push_back(rt[f]);
            for (int a = 0; a < npf; ++a) {
                const int32_t r = pb->h_t2a[pb->h_ien[e * npe + fn[f][a]]];
                if (r >= 0) rows[r].push_back(s * npf + a);
            }
        }
    }
    RSD_CHECK((int64_t)faces.size() * npf < INT32_MAX, "natural boundary conditions: too many tagged faces");
    R.n_faces = (int64_t)faces.size();
    std::vector<int32_t> hrow, hoff(1, 0), hidx;
    for (auto& [r, list] : rows) {
        hrow.push_back(r);
        hidx.insert(hidx.end(), list.begin(), list.end());
        hoff.push_back((int32_t)hidx.size());
    }
    R.n_rows = (int64_t)hrow.size();
    hipStream_t st = pb->ctx->stream;
    R.face.upload(faces.data(), R.n_faces, st);
    R.ftag.upload(ftags.data(), R.n_faces, st);
    R.row.upload(hrow.data(), R.n_rows, st);
    R.off.upload(hoff.data(), (int64_t)hoff.size(), st);
    R.idx.upload(hidx.data(), (int64_t)hidx.size(), st);
    R.fe.alloc(R.n_faces * npf);
    R.ex.alloc(kNatMaxEntries * 3);
    RSD_HIP(hipStreamSynchronize(st));  // the host vectors above go out of scope
    slot = plan;
    return R;
}

// One thread per tagged face s: Fe of face f of element e, every point's
// entries in order, into fe[s * NPF ..].
template<int D, int P, int NQ>
__global__ __launch_bounds__(kBlock) void k_natural_faces(AsmArgs A, NatEntries E, const int32_t* __restrict__ face,
                                                           const int32_t* __restrict__ ftag, int64_t n_faces,
                                                           double* __restrict__ fe) {
    constexpr int NPE = Elem<D, P>::NPE, NF = 2 * D, FQ = D == 2 ? NQ : NQ * NQ;
    constexpr int NPF = D == 2 ? P + 1 : (P + 1) * (P + 1);
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= n_faces) return;
    const int32_t ef = face[s];
    const int f = ef % NF;
    const int64_t e = ef / NF;
    const int32_t tag = ftag[s];
    int32_t nodes[NPE];
    double X[NPE][D];
    load_element<D, P>(A, e, nodes, X);
    // local face nodes (face_local_nodes: in-face axes ascending, lower fastest)
    int fl[NPF];
    {
        const int n1 = P + 1, axis = f / 2, fixed = (f & 1) ? P : 0;
        int fr0 = axis == 0 ? 1 : 0, fr1 = axis == 2 ? 1 : 2;
        int k = 0;
        const int outer = D == 3 ? P : 0;
        for (int v = 0; v <= outer; ++v)
            for (int u = 0; u <= P; ++u) {
                int c[3] = {0, 0, 0};
                c[axis] = fixed;
                c[fr0] = u;
                if (D == 3) c[fr1] = v;
                fl[k++] = c[0] + n1 * c[1] + n1 * n1 * c[2];
            }
    }
    double Fe[NPF];
#pragma unroll
    for (int a = 0; a < NPF; ++a) Fe[a] = 0.0;
    for (int q = 0; q < FQ; ++q) {
        double xi[3], nref[D];
        const double w = face_point<D, NQ>(f, q, xi, nref);
        QP<D, NPE> Q;
        double J[D * D];
        qp_at<D, P, false>(xi, X, Q, true, J);
        double n[D];
        boundary_normal<D>(J, nref, n);
        double Nf[NPF];
#pragma unroll
        for (int a = 0; a < NPF; ++a) Nf[a] = Q.N[fl[a]];
        for (int k = 0; k < E.n; ++k) {
            if (E.tag[k] != tag) continue;
            double val[D];
            if (E.nodal[k]) {
                // NodalFluxForm: val = sum_b Nface_b q(node_b), from 0
#pragma unroll
                for (int i = 0; i < D; ++i) val[i] = 0.0;
#pragma unroll
                for (int b = 0; b < NPF; ++b)
#pragma unroll
                    for (int i = 0; i < D; ++i) val[i] += Nf[b] * E.nodal[k][(int64_t)nodes[fl[b]] * D + i];
            } else {
                const double z = D == 3 ? Q.x[D - 1] : 0.0;
#pragma unroll
                for (int i = 0; i < D; ++i) val[i] = eval_expr(E.ex[k * D + i], A.time, Q.x[0], Q.x[1], z, 0.0);
            }
#pragma unroll
            for (int a = 0; a < NPF; ++a)
#pragma unroll
                for (int i = 0; i < D; ++i) Fe[a] += ((val[i] * Nf[a]) * n[i]) * w;
        }
    }
#pragma unroll
    for (int a = 0; a < NPF; ++a) fe[s * NPF + a] = Fe[a];
}

// F[row[i]] += its terms, one after another in face order
__global__ void k_rows_add(int64_t n_rows, const int32_t* __restrict__ row, const int32_t* __restrict__ off,
                           const int32_t* __restrict__ idx, const double* __restrict__ fe, double* __restrict__ F) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_rows) return;
    const int32_t r = row[i];
    double acc = F[r];
    for (int32_t j = off[i]; j < off[i + 1]; ++j) acc += 1.0 * fe[idx[j]];
    F[r] = acc;
}

template<int D, int P, int NQ>
static void run_natural(rsd_problem* pb, const AsmArgs& A, const NatEntries& E, NaturalPlan& R, double* F) {
    hipStream_t st = pb->ctx->stream;
    if (R.n_faces)
        hipLaunchKernelGGL((k_natural_faces<D, P, NQ>), dim3((unsigned)nblocks(R.n_faces)), dim3(kBlock), 0, st, A, E,
                           R.face.p, R.ftag.p, R.n_faces, R.fe.p);
    if (R.n_rows)
        hipLaunchKernelGGL(k_rows_add, dim3((unsigned)nblocks(R.n_rows)), dim3(kBlock), 0, st, R.n_rows, R.row.p,
                           R.off.p, R.idx.p, R.fe.p, F);
    RSD_HIP(hipGetLastError());
}

void apply_natural(rsd_problem* pb, double time, int32_t n_entries, const int32_t* tags, const char* const* exprs,
                   const double* const* nodal, double* F) {
    RSD_CHECK(n_entries >= 0 && n_entries <= kNatMaxEntries, "natural boundary conditions: 0 to 8 entries");
    if (n_entries == 0) return;
    const int D = pb->dim;
    NatEntries E{};
    E.n = n_entries;
    std::vector<Expr> hex((size_t)kNatMaxEntries * 3);
    std::vector<int32_t> distinct;
    for (int k = 0; k < n_entries; ++k) {
        E.tag[k] = tags[k];
        E.nodal[k] = nodal ? nodal[k] : nullptr;
        if (!E.nodal[k]) {
            RSD_CHECK(exprs, "natural boundary conditions: an entry has neither expressions nor nodal values");
            for (int i = 0; i < D; ++i) {
                RSD_CHECK(exprs[k * D + i], "natural boundary conditions: missing flux expression");
                hex[(size_t)k * D + i] = compile_expr(exprs[k * D + i]);
                RSD_CHECK(!hex[(size_t)k * D + i].uses_T, "natural boundary conditions: a flux may not depend on T");
            }
        }
        if (std::find(distinct.begin(), distinct.end(), tags[k]) == distinct.end()) distinct.push_back(tags[k]);
    }
    std::sort(distinct.begin(), distinct.end());
    NaturalPlan& R = natural_plan(pb, distinct);
    RSD_HIP(hipMemcpyAsync(R.ex.p, hex.data(), sizeof(Expr) * hex.size(), hipMemcpyHostToDevice, pb->ctx->stream));
    RSD_HIP(hipStreamSynchronize(pb->ctx->stream));  // hex goes out of scope
    E.ex = R.ex.p;
    AsmArgs A = base_args(pb);
    A.time = time;
    if (pb->dim == 2 && pb->p == 1 && pb->nq == 2) { run_natural<2, 1, 2>(pb, A, E, R, F); return; }
    if (pb->dim == 2 && pb->p == 1 && pb->nq == 3) { run_natural<2, 1, 3>(pb, A, E, R, F); return; }
    if (pb->dim == 2 && pb->p == 2 && pb->nq == 3) { run_natural<2, 2, 3>(pb, A, E, R, F); return; }
    if (pb->dim == 2 && pb->p == 2 && pb->nq == 4) { run_natural<2, 2, 4>(pb, A, E, R, F); return; }
    if (pb->dim == 3 && pb->p == 1 && pb->nq == 2) { run_natural<3, 1, 2>(pb, A, E, R, F); return; }
    if (pb->dim == 3 && pb->p == 1 && pb->nq == 3) { run_natural<3, 1, 3>(pb, A, E, R, F); return; }
    throw Error(RSD_ERR_ARG, "natural boundary conditions: discretisation not supported (dim/p/nq)");
}

}  // namespace rsd

using namespace rsd;

extern "C" int rsd_apply_natural(rsd_ctx* ctx, rsd_problem* pb, double time, int32_t n_entries,
                                 const int32_t* entry_tags, const char* const* flux_exprs,
                                 const double* const* d_nodal_flux, double* F) {
    return guard([&] {
        RSD_CHECK(ctx && pb && F && (n_entries == 0 || entry_tags), "rsd_apply_natural: null argument");
        RSD_CHECK(ctx == pb->ctx, "problem belongs to another context");
        apply_natural(pb, time, n_entries, entry_tags, flux_exprs, d_nodal_flux, F);
    });
}
