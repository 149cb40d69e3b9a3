// rsd_face.hpp -- element faces: local face nodes, face quadrature points
// and the unnormalised boundary normal, shared by the boundary monitors and
// the natural boundary conditions.
#pragma once

#include "rsd_element.hpp"

#include <vector>

namespace rsd {

// local nodes of face f in the reference's order (LagrangeQuad.tpp:484-503,
// LagrangeHex.tpp:263-304): the two in-face axes ascending, the lower one
// fastest
inline std::vector<int32_t> face_local_nodes(int D, int P, int f) {
    const int n1 = P + 1, axis = f / 2, fixed = (f & 1) ? P : 0;
    int fr[2] = {0, 0}, m = 0;
    for (int ax = 0; ax < D; ++ax)
        if (ax != axis) fr[m++] = ax;
    std::vector<int32_t> out;
    const int outer = D == 3 ? P : 0;
    for (int v = 0; v <= outer; ++v)
        for (int u = 0; u <= P; ++u) {
            int c[3] = {0, 0, 0};
            c[axis] = fixed;
            c[fr[0]] = u;
            if (D == 3) c[fr[1]] = v;
            out.push_back(c[0] + n1 * c[1] + n1 * n1 * c[2]);
        }
    return out;
}

// face point q of face f: element coordinates and weight (the face rule is
// the 1D / tensor Gauss rule, first in-face axis fastest), and the face's
// reference normal (mapFaceToElement / getFaceTopology, LagrangeQuad.tpp:
// 95-181, LagrangeHex.tpp:120-267)
template<int D, int NQ>
__device__ __forceinline__ double face_point(int f, int q, double (&xi)[3], double (&nref)[D]) {
    const double side = (f & 1) ? 1.0 : -1.0;
    double fw;
    xi[2] = 0.0;
    if constexpr (D == 2) {
        double s;
        gauss<NQ>(q, s, fw);
        xi[0] = f < 2 ? side : s;
        xi[1] = f < 2 ? s : side;
    } else {
        double s, t, ws, wt;
        gauss<NQ>(q % NQ, s, ws);
        gauss<NQ>(q / NQ, t, wt);
        fw = ws * wt;
        if (f < 2) { xi[0] = side; xi[1] = s; xi[2] = t; }
        else if (f < 4) { xi[0] = s; xi[1] = side; xi[2] = t; }
        else { xi[0] = s; xi[1] = t; xi[2] = side; }
    }
#pragma unroll
    for (int i = 0; i < D; ++i) nref[i] = i == f / 2 ? side : 0.0;
    return fw;
}

// the unnormalised boundary normal cof(J) n_ref (JacobianTransform::
// computeBoundaryNormal with computeMatrixCofactor, JacobianTransform.tpp:
// 110-123, 209-233)
template<int D>
__device__ __forceinline__ void boundary_normal(const double* A, const double (&nref)[D], double (&n)[D]) {
    double c[D * D];
    if constexpr (D == 2) {
        c[0] = A[3]; c[1] = -A[2]; c[2] = -A[1]; c[3] = A[0];
    } else {
        c[0] = (A[4] * A[8] - A[5] * A[7]);
        c[1] = -(A[3] * A[8] - A[5] * A[6]);
        c[2] = (A[3] * A[7] - A[4] * A[6]);
        c[3] = -(A[1] * A[8] - A[2] * A[7]);
        c[4] = (A[0] * A[8] - A[2] * A[6]);
        c[5] = -(A[0] * A[7] - A[1] * A[6]);
        c[6] = (A[1] * A[5] - A[2] * A[4]);
        c[7] = -(A[0] * A[5] - A[2] * A[3]);
        c[8] = (A[0] * A[4] - A[1] * A[3]);
    }
#pragma unroll
    for (int i = 0; i < D; ++i) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) s += c[i * D + j] * nref[j];
        n[i] = s;
    }
}

}  // namespace rsd
