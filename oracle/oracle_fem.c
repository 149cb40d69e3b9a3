/*
 * oracle_fem.c -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Serial restatement of the reference's mesh generation, constraint
 * numbering, CSR pattern construction, element math and assembly loops.
 * Compiled with -ffp-contract=off: every product and sum is rounded on its
 * own, in the operand order the reference's C++ evaluates them.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXNPE 64  /* (kMaxBasisOrder + 1)^3, DiscretizationLimits.hpp */
#define MAXQ1D 5

/* ------------------------------------------------------------------ */
/* mesh: src/mesh/generator/BlockMesh2D.cpp:50-121, BlockMesh3D.cpp:50-139 */
/* ------------------------------------------------------------------ */
void orc_block_mesh_2d(int64_t nx, int64_t ny, double x0, double x1, double y0, double y1,
                       int64_t px, int64_t py, double* xyz, int64_t* ien, int32_t* rng) {
    const int64_t NX = nx * px + 1, NY = ny * py + 1;
    const double dx = (x1 - x0) / (double)(nx * px);
    const double dy = (y1 - y0) / (double)(ny * py);
    int64_t node = 0;
    for (int64_t j = 0; j < NY; ++j)
        for (int64_t i = 0; i < NX; ++i) {
            xyz[2 * node] = x0 + (double)i * dx;
            xyz[2 * node + 1] = y0 + (double)j * dy;
            node++;
        }
    const int64_t npe = (px + 1) * (py + 1);
    for (int64_t ey = 0; ey < ny; ++ey)
        for (int64_t ex = 0; ex < nx; ++ex) {
            const int64_t e = ey * nx + ex;
            for (int64_t a = 0; a < npe; ++a) {
                const int64_t ax = a % (px + 1), ay = a / (px + 1);
                ien[e * npe + a] = (ey * py + ay) * (nx * px + 1) + (ex * px + ax);
            }
            rng[e * 4 + 0] = (ex == 0) ? 0 : -1;
            rng[e * 4 + 1] = (ex == nx - 1) ? 1 : -1;
            rng[e * 4 + 2] = (ey == 0) ? 2 : -1;
            rng[e * 4 + 3] = (ey == ny - 1) ? 3 : -1;
        }
}

void orc_block_mesh_3d(int64_t nx, int64_t ny, int64_t nz, double x0, double x1, double y0,
                       double y1, double z0, double z1, int64_t px, int64_t py, int64_t pz,
                       double* xyz, int64_t* ien, int32_t* rng) {
    const int64_t NX = nx * px + 1, NY = ny * py + 1, NZ = nz * pz + 1;
    const double dx = (x1 - x0) / (double)(nx * px);
    const double dy = (y1 - y0) / (double)(ny * py);
    const double dz = (z1 - z0) / (double)(nz * pz);
    int64_t node = 0;
    for (int64_t k = 0; k < NZ; ++k)
        for (int64_t j = 0; j < NY; ++j)
            for (int64_t i = 0; i < NX; ++i) {
                xyz[3 * node] = x0 + (double)i * dx;
                xyz[3 * node + 1] = y0 + (double)j * dy;
                xyz[3 * node + 2] = z0 + (double)k * dz;
                node++;
            }
    const int64_t npe = (px + 1) * (py + 1) * (pz + 1);
    for (int64_t ez = 0; ez < nz; ++ez)
        for (int64_t ey = 0; ey < ny; ++ey)
            for (int64_t ex = 0; ex < nx; ++ex) {
                const int64_t e = ez * (nx * ny) + ey * nx + ex;
                for (int64_t a = 0; a < npe; ++a) {
                    const int64_t ax = a % (px + 1);
                    const int64_t ay = (a / (px + 1)) % (py + 1);
                    const int64_t az = a / ((px + 1) * (py + 1));
                    ien[e * npe + a] = (ez * pz + az) * (NY * NX) + (ey * py + ay) * NX + (ex * px + ax);
                }
                rng[e * 6 + 0] = (ex == 0) ? 0 : -1;
                rng[e * 6 + 1] = (ex == nx - 1) ? 1 : -1;
                rng[e * 6 + 2] = (ey == 0) ? 2 : -1;
                rng[e * 6 + 3] = (ey == ny - 1) ? 3 : -1;
                rng[e * 6 + 4] = (ez == 0) ? 4 : -1;
                rng[e * 6 + 5] = (ez == nz - 1) ? 5 : -1;
            }
}

/* face -> local node list: LagrangeQuad.tpp:484-503, LagrangeHex.tpp:263-304 */
static int face_nodes(int dim, const int32_t* p, int f, int64_t* out) {
    const int64_t X = p[0], Y = p[1];
    int n = 0;
    if (dim == 2) {
        switch (f) {
        case 0: for (int64_t i = 0; i <= Y; ++i) out[n++] = i * (X + 1); break;
        case 1: for (int64_t i = 0; i <= Y; ++i) out[n++] = i * (X + 1) + X; break;
        case 2: for (int64_t i = 0; i <= X; ++i) out[n++] = i; break;
        case 3: for (int64_t i = 0; i <= X; ++i) out[n++] = i + (X + 1) * Y; break;
        }
        return n;
    }
    const int64_t Z = p[2];
    switch (f) {
    case 0: for (int64_t j = 0; j <= Z; ++j) for (int64_t i = 0; i <= Y; ++i) out[n++] = i * (X + 1) + j * (X + 1) * (Y + 1); break;
    case 1: for (int64_t j = 0; j <= Z; ++j) for (int64_t i = 0; i <= Y; ++i) out[n++] = i * (X + 1) + j * (X + 1) * (Y + 1) + X; break;
    case 2: for (int64_t j = 0; j <= Z; ++j) for (int64_t i = 0; i <= X; ++i) out[n++] = i + j * (X + 1) * (Y + 1); break;
    case 3: for (int64_t j = 0; j <= Z; ++j) for (int64_t i = 0; i <= X; ++i) out[n++] = i + j * (X + 1) * (Y + 1) + (X + 1) * Y; break;
    case 4: for (int64_t j = 0; j <= Y; ++j) for (int64_t i = 0; i <= X; ++i) out[n++] = i + j * (X + 1); break;
    case 5: for (int64_t j = 0; j <= Y; ++j) for (int64_t i = 0; i <= X; ++i) out[n++] = i + j * (X + 1) + (X + 1) * (Y + 1) * Z; break;
    }
    return n;
}

/* TopologicalDOF::buildConstraints, topology/TopologicalDOF.tpp:111-190
 * (one DOF per node, interleaved ordering).  The constraint tag of a node
 * is the last essential tag that touched it in (element, face) order. */
int64_t orc_build_constraints(int32_t dim, const int32_t* p, int64_t n_nodes, int64_t n_elem,
                              const int64_t* ien, const int32_t* rng, int32_t n_tags,
                              const int32_t* tags, int32_t* topo_to_alg, int32_t* alg_to_topo,
                              int32_t* ctag) {
    const int nfaces = 2 * dim;
    int64_t npe = (int64_t)(p[0] + 1) * (p[1] + 1) * (dim == 3 ? (p[2] + 1) : 1);
    char* constrained = (char*)calloc((size_t)n_nodes, 1);
    for (int64_t i = 0; i < n_nodes; ++i) ctag[i] = -1;
    int64_t fn[64];
    for (int64_t e = 0; e < n_elem; ++e)
        for (int f = 0; f < nfaces; ++f) {
            const int32_t tag = rng[e * nfaces + f];
            if (tag < 0) continue;
            int essential = 0;
            for (int k = 0; k < n_tags; ++k) if (tags[k] == tag) essential = 1;
            const int npf = face_nodes(dim, p, f, fn);
            for (int i = 0; i < npf; ++i) {
                if (!essential) continue;
                const int64_t g = ien[e * npe + fn[i]];
                constrained[g] = 1;
                ctag[g] = tag;
            }
        }
    int64_t alg = 0;
    for (int64_t i = 0; i < n_nodes; ++i) {
        if (constrained[i]) { topo_to_alg[i] = -1; continue; }
        topo_to_alg[i] = (int32_t)alg;
        if (alg_to_topo) alg_to_topo[alg] = (int32_t)i;
        ++alg;
    }
    free(constrained);
    return alg;
}

/* ------------------------------------------------------------------ */
/* CSR pattern: Assembler.tpp:7-76                                      */
/* ------------------------------------------------------------------ */
static int cmp_i64(const void* a, const void* b) {
    int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return (x > y) - (x < y);
}

int64_t orc_create_matrix(int64_t n_elem, int32_t npe, const int64_t* ien,
                          const int32_t* topo_to_alg, int64_t n_free, int64_t* row_ptr,
                          int64_t* col_idx) {
    /* pass 1: raw touches per row */
    int64_t* off = (int64_t*)calloc((size_t)(n_free + 1), sizeof(int64_t));
    for (int64_t e = 0; e < n_elem; ++e)
        for (int i = 0; i < npe; ++i) {
            const int32_t r = topo_to_alg[ien[e * npe + i]];
            if (r < 0) continue;
            for (int k = 0; k < npe; ++k)
                if (topo_to_alg[ien[e * npe + k]] >= 0) off[r + 1]++;
        }
    for (int64_t i = 0; i < n_free; ++i) off[i + 1] += off[i];
    /* pass 2: fill via cursors */
    int64_t* raw = (int64_t*)malloc(sizeof(int64_t) * (size_t)(off[n_free] ? off[n_free] : 1));
    int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n_free ? n_free : 1));
    for (int64_t i = 0; i < n_free; ++i) cur[i] = off[i];
    for (int64_t e = 0; e < n_elem; ++e)
        for (int i = 0; i < npe; ++i) {
            const int32_t r = topo_to_alg[ien[e * npe + i]];
            if (r < 0) continue;
            for (int k = 0; k < npe; ++k) {
                const int32_t c = topo_to_alg[ien[e * npe + k]];
                if (c >= 0) raw[cur[r]++] = c;
            }
        }
    /* sort + unique per row */
    row_ptr[0] = 0;
    for (int64_t i = 0; i < n_free; ++i) {
        int64_t* b = raw + off[i];
        const int64_t cnt = off[i + 1] - off[i];
        qsort(b, (size_t)cnt, sizeof(int64_t), cmp_i64);
        int64_t u = 0;
        for (int64_t k = 0; k < cnt; ++k)
            if (u == 0 || b[k] != b[u - 1]) b[u++] = b[k];
        row_ptr[i + 1] = row_ptr[i] + u;
        if (col_idx) memcpy(col_idx + row_ptr[i], b, sizeof(int64_t) * (size_t)u);
    }
    const int64_t nnz = row_ptr[n_free];
    free(off); free(raw); free(cur);
    return nnz;
}

/* CSRMatrix::getDataIndex, linalg/types/CSRMatrix.hpp:49-73 */
static int64_t data_index(const int64_t* row_ptr, const int64_t* col_idx, int64_t i, int64_t j) {
    int64_t s = row_ptr[i], e = row_ptr[i + 1];
    while (s < e) {
        const int64_t m = s + (e - s) / 2;
        if (col_idx[m] == j) return m;
        if (col_idx[m] < j) s = m + 1; else e = m;
    }
    return -1;
}

void orc_elem_nnz_map(int64_t n_elem, int32_t npe, const int64_t* ien, const int32_t* topo_to_alg,
                      const int64_t* row_ptr, const int64_t* col_idx, int64_t* map) {
    for (int64_t e = 0; e < n_elem; ++e)
        for (int a = 0; a < npe; ++a) {
            const int32_t r = topo_to_alg[ien[e * npe + a]];
            for (int b = 0; b < npe; ++b) {
                const int32_t c = topo_to_alg[ien[e * npe + b]];
                map[(e * npe + a) * npe + b] = (r < 0 || c < 0) ? -1 : data_index(row_ptr, col_idx, r, c);
            }
        }
}

int32_t orc_colour_order(int64_t n_elem, int32_t npe, const int64_t* ien, int64_t n_nodes,
                         int32_t* colour, int64_t* order) {
    uint64_t* mask = (uint64_t*)calloc((size_t)n_nodes, sizeof(uint64_t));
    int32_t nc = 0;
    for (int64_t e = 0; e < n_elem; ++e) {
        uint64_t used = 0;
        for (int a = 0; a < npe; ++a) used |= mask[ien[e * npe + a]];
        int c = 0;
        while (used & (1ull << c)) ++c;
        colour[e] = c;
        if (c + 1 > nc) nc = c + 1;
        for (int a = 0; a < npe; ++a) mask[ien[e * npe + a]] |= (1ull << c);
    }
    int64_t k = 0;
    for (int c = 0; c < nc; ++c)
        for (int64_t e = 0; e < n_elem; ++e)
            if (colour[e] == c) order[k++] = e;
    free(mask);
    return nc;
}

/* ------------------------------------------------------------------ */
/* element math                                                         */
/* ------------------------------------------------------------------ */
/* Lagrange1D::eval / evalFirstDerivative, fem/basis/Lagrange1D.tpp:4-51 */
static void lag1d(int p, double xi, double* N, double* dN) {
    if (p == 1) {
        N[0] = 0.5 * (1.0 - xi); N[1] = 0.5 * (1.0 + xi);
        dN[0] = -0.5; dN[1] = 0.5;
    } else if (p == 2) {
        N[0] = 0.5 * xi * (xi - 1.0); N[1] = (1.0 - xi) * (1.0 + xi); N[2] = 0.5 * xi * (xi + 1.0);
        dN[0] = xi - 0.5; dN[1] = -2.0 * xi; dN[2] = xi + 0.5;
    } else {
        N[0] = (-9.0 / 16.0) * (xi - (1.0 / 3.0)) * (xi - 1.0) * (xi + (1.0 / 3.0));
        N[1] = (27.0 / 16.0) * (xi - 1.0) * (xi - (1.0 / 3.0)) * (xi + 1.0);
        N[2] = (-27.0 / 16.0) * (xi + (1.0 / 3.0)) * (xi - 1.0) * (xi + 1.0);
        N[3] = (9.0 / 16.0) * (xi - (1.0 / 3.0)) * (xi + 1.0) * (xi + (1.0 / 3.0));
        dN[0] = (-9.0 / 16.0) * (3.0 * xi * xi - 2.0 * xi - (1.0 / 9.0));
        dN[1] = (27.0 / 16.0) * (3.0 * xi * xi - (2.0 / 3.0) * xi - 1.0);
        dN[2] = (-27.0 / 16.0) * (3.0 * xi * xi + (2.0 / 3.0) * xi - 1.0);
        dN[3] = (9.0 / 16.0) * (3.0 * xi * xi + 2.0 * xi - (1.0 / 9.0));
    }
}

/* GaussQuadrature1D, fem/quadrature/GaussQuadrature1D.tpp:4-71 */
static void gauss1d(int n, double* x, double* w) {
    static const double X[6][5] = {
        {0}, {0.0},
        {-0.5773502691896257, 0.5773502691896257},
        {-0.7745966692414834, 0.0, 0.7745966692414834},
        {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526},
        {-0.9061798459386640, -0.5384693101056831, 0.0, 0.5384693101056831, 0.9061798459386640}};
    static const double W[6][5] = {
        {0}, {2.0},
        {1.0, 1.0},
        {0.5555555555555556, 0.8888888888888889, 0.5555555555555556},
        {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538},
        {0.2369268850561891, 0.4786286704993665, 0.5688888888888889, 0.4786286704993665, 0.2369268850561891}};
    for (int i = 0; i < n; ++i) { x[i] = X[n][i]; w[i] = W[n][i]; }
}

/* tensor-product points/weights: GaussQuadratureQuad.tpp:4-40,
 * GaussQuadratureHex.tpp:4-49 (x fastest) */
static int quad_points(const orc_problem* P, double* xi, double* w) {
    double x[3][MAXQ1D], ww[3][MAXQ1D];
    for (int d = 0; d < P->dim; ++d) gauss1d(P->nq[d], x[d], ww[d]);
    int q = 0;
    if (P->dim == 2) {
        for (int j = 0; j < P->nq[1]; ++j)
            for (int i = 0; i < P->nq[0]; ++i) {
                xi[2 * q] = x[0][i]; xi[2 * q + 1] = x[1][j];
                w[q] = ww[0][i] * ww[1][j]; q++;
            }
    } else {
        for (int k = 0; k < P->nq[2]; ++k)
            for (int j = 0; j < P->nq[1]; ++j)
                for (int i = 0; i < P->nq[0]; ++i) {
                    xi[3 * q] = x[0][i]; xi[3 * q + 1] = x[1][j]; xi[3 * q + 2] = x[2][k];
                    w[q] = ww[0][i] * ww[1][j] * ww[2][k]; q++;
                }
    }
    return q;
}

typedef struct {
    double N[MAXNPE], dNdxi[3 * MAXNPE], dNdx[3 * MAXNPE];
    double x[3], J[9], g[9], measure, w;
    double K[9], dKdT[9], rho, cp, dcpdT, T, Trate, gradT[3];
} qp_t;

static double det(int d, const double* A) {
    if (d == 2) return (A[0] * A[3]) - (A[1] * A[2]);
    return (A[0] * ((A[4] * A[8]) - (A[5] * A[7])))
         - (A[1] * ((A[3] * A[8]) - (A[5] * A[6])))
         + (A[2] * ((A[3] * A[7]) - (A[4] * A[6])));
}

/* EvalQuadraturePointVolume::evaluate (EvalQuadraturePointVolume.hpp:64-85)
 * with LagrangeQuad/Hex eval+evalGradient (LagrangeQuad.tpp:4-44,
 * LagrangeHex.tpp:4-55) and JacobianTransform (JacobianTransform.tpp:3-107,
 * 159-206), for SpatialDim == ParametricDim. */
static void qp_evaluate(const orc_problem* P, const double* coords, const double* xi, double w, qp_t* q) {
    const int d = P->dim, npe = P->npe;
    double Nx[3][4], dNx[3][4];
    for (int k = 0; k < d; ++k) lag1d(P->p[k], xi[k], Nx[k], dNx[k]);
    int a = 0;
    if (d == 2) {
        for (int j = 0; j <= P->p[1]; ++j)
            for (int i = 0; i <= P->p[0]; ++i) {
                q->N[a] = Nx[0][i] * Nx[1][j];
                q->dNdxi[a * 2] = dNx[0][i] * Nx[1][j];
                q->dNdxi[a * 2 + 1] = Nx[0][i] * dNx[1][j];
                a++;
            }
    } else {
        for (int k = 0; k <= P->p[2]; ++k)
            for (int j = 0; j <= P->p[1]; ++j)
                for (int i = 0; i <= P->p[0]; ++i) {
                    q->N[a] = Nx[0][i] * Nx[1][j] * Nx[2][k];
                    q->dNdxi[a * 3] = dNx[0][i] * Nx[1][j] * Nx[2][k];
                    q->dNdxi[a * 3 + 1] = Nx[0][i] * dNx[1][j] * Nx[2][k];
                    q->dNdxi[a * 3 + 2] = Nx[0][i] * Nx[1][j] * dNx[2][k];
                    a++;
                }
    }
    q->w = w;
    for (int i = 0; i < d; ++i) q->x[i] = 0.0;
    for (int b = 0; b < npe; ++b)
        for (int i = 0; i < d; ++i) q->x[i] += q->N[b] * coords[b * d + i];
    for (int i = 0; i < d * d; ++i) q->J[i] = 0.0;
    for (int i = 0; i < d; ++i)
        for (int al = 0; al < d; ++al)
            for (int b = 0; b < npe; ++b)
                q->J[i * d + al] += q->dNdxi[b * d + al] * coords[b * d + i];
    for (int be = 0; be < d; ++be)
        for (int al = 0; al < d; ++al) {
            double s = 0.0;
            for (int i = 0; i < d; ++i) s += q->J[i * d + al] * q->J[i * d + be];
            q->g[be * d + al] = s;
            q->g[al * d + be] = s;
        }
    q->measure = sqrt(det(d, q->g));
    double inv[9];
    const double A0 = det(d, q->J);
    const double di = 1.0 / A0;
    const double* A = q->J;
    if (d == 2) {
        inv[0] = di * A[3]; inv[1] = -di * A[1]; inv[2] = -di * A[2]; inv[3] = di * A[0];
    } else {
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
    for (int i = 0; i < npe * d; ++i) q->dNdx[i] = 0.0;
    for (int b = 0; b < npe; ++b)
        for (int i = 0; i < d; ++i)
            for (int al = 0; al < d; ++al)
                q->dNdx[b * d + i] += inv[i * d + al] * q->dNdxi[b * d + al];
}

/* HeatStateVector::interpolate (HeatStateVector.hpp:51-56) via
 * TemperatureField::eval/evalGradient (TemperatureField.hpp:14-39) */
static void qp_interpolate(const orc_problem* P, const double* Uval, const double* Urate, qp_t* q) {
    const int d = P->dim, npe = P->npe;
    double u = 0;
    for (int a = 0; a < npe; ++a) u += q->N[a] * Uval[a];
    q->T = u;
    for (int i = 0; i < d; ++i) q->gradT[i] = 0;
    for (int a = 0; a < npe; ++a)
        for (int i = 0; i < d; ++i) q->gradT[i] += q->dNdx[a * d + i] * Uval[a];
    double r = 0;
    for (int a = 0; a < npe; ++a) r += q->N[a] * Urate[a];
    q->Trate = r;
}

/* ConductivityModel::eval/evalGradient (ConductivityModel.hpp:60-96),
 * DensityModel (DensityModel.hpp:13), SpecificHeatModel (:35-61); the
 * per-element factor kappa_e multiplies k(T) and k'(T). */
static void qp_models(const orc_problem* P, int64_t e, qp_t* q) {
    const int d = P->dim;
    double k = P->k_scalar, dk = 0.0;
    if (P->k_expr) k = orc_expr_eval(P->k_expr, 0, 0, 0, 0, q->T);
    if (P->dk_expr) dk = orc_expr_eval(P->dk_expr, 0, 0, 0, 0, q->T);
    if (P->kappa_e) { k = P->kappa_e[e] * k; dk = P->kappa_e[e] * dk; }
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) {
            const double shape = P->anisotropic ? P->k_tensor[i * d + j] : ((i == j) ? 1.0 : 0.0);
            q->K[i * d + j] = k * shape;
        }
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) {
            const double shape = P->anisotropic ? P->k_tensor[i * d + j] : ((i == j) ? 1.0 : 0.0);
            q->dKdT[i * d + j] = dk * shape;
        }
    q->rho = P->rho;
    q->cp = P->cp_expr ? orc_expr_eval(P->cp_expr, 0, 0, 0, 0, q->T) : P->cp;
    q->dcpdT = P->dcp_expr ? orc_expr_eval(P->dcp_expr, 0, 0, 0, 0, q->T) : 0.0;
}

/* one QP's contribution of one form to Ke: DiffusionForm.hpp:11-31,
 * MassForm.hpp:11-16, TangentDiffusionForm.hpp:11-31, TangentMassForm.hpp:11-16;
 * a ScaledForm (ScaledForm.hpp:14-31) adds coef * contribution. */
static void form_matrix(const orc_problem* P, int kind_flags, double coef, const qp_t* q, double* Ke) {
    const int d = P->dim, n = P->npe;
    const int kind = kind_flags & 0xff, scaled = (kind_flags & ORC_SCALED) != 0;
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            double c;
            if (kind == ORC_FORM_DIFFUSION || kind == ORC_FORM_TDIFFUSION) {
                double integrand = 0.0;
                for (int i = 0; i < d; ++i) {
                    double mv = 0.0;
                    if (kind == ORC_FORM_DIFFUSION)
                        for (int j = 0; j < d; ++j) mv += q->K[i * d + j] * q->dNdx[b * d + j];
                    else
                        for (int j = 0; j < d; ++j) mv += q->dKdT[i * d + j] * q->gradT[j];
                    integrand += q->dNdx[a * d + i] * mv;
                }
                c = (kind == ORC_FORM_DIFFUSION) ? integrand * q->measure * q->w
                                                 : integrand * q->N[b] * q->measure * q->w;
            } else if (kind == ORC_FORM_MASS) {
                c = q->rho * q->cp * q->N[a] * q->N[b] * q->measure * q->w;
            } else {
                c = q->rho * q->dcpdT * q->Trate * q->N[a] * q->N[b] * q->measure * q->w;
            }
            if (scaled) Ke[a * n + b] += coef * c;
            else Ke[a * n + b] += c;
        }
}

/* DiffusionForm.hpp:33-52, MassForm.hpp:18-27, TangentDiffusionForm.hpp:33-56,
 * TangentMassForm.hpp:18-27 (vector forms, per-QP contribution) */
static void form_vector(const orc_problem* P, int kind_flags, double coef, const qp_t* q, const double* Ue, double* Fe) {
    const int d = P->dim, n = P->npe;
    const int kind = kind_flags & 0xff, scaled = (kind_flags & ORC_SCALED) != 0;
    double sbuf[MAXNPE];
    /* an unscaled form accumulates straight into Fe; a ScaledForm fills a
     * zeroed buffer and then adds coef * buffer (ScaledForm.hpp:34-47) */
    double* buf = scaled ? sbuf : Fe;
    if (scaled) for (int a = 0; a < n; ++a) buf[a] = 0.0;
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            if (kind == ORC_FORM_DIFFUSION || kind == ORC_FORM_TDIFFUSION) {
                double integrand = 0.0;
                for (int i = 0; i < d; ++i) {
                    double mv = 0.0;
                    if (kind == ORC_FORM_DIFFUSION)
                        for (int j = 0; j < d; ++j) mv += q->K[i * d + j] * q->dNdx[b * d + j];
                    else
                        for (int j = 0; j < d; ++j) mv += q->dKdT[i * d + j] * q->gradT[j];
                    integrand += q->dNdx[a * d + i] * mv;
                }
                if (kind == ORC_FORM_DIFFUSION) buf[a] += integrand * Ue[b] * q->measure * q->w;
                else buf[a] += integrand * q->N[b] * Ue[b] * q->measure * q->w;
            } else if (kind == ORC_FORM_MASS) {
                buf[a] += q->rho * q->cp * q->N[a] * q->N[b] * Ue[b] * q->measure * q->w;
            } else {
                buf[a] += q->rho * q->dcpdT * q->Trate * q->N[a] * q->N[b] * Ue[b] * q->measure * q->w;
            }
        }
    if (scaled) for (int a = 0; a < n; ++a) Fe[a] += coef * buf[a];
}

/* gatherElementVector, ElementMap.tpp:18-75.  mode 0 Free, 1 Full, 2 Constrained */
static double bc_value(const orc_problem* P, int32_t tag, double time, const double* x) {
    double v = 0.0;
    for (int k = 0; k < P->n_bc; ++k)
        if (P->bc_tag[k] == tag)
            v = orc_expr_eval(P->bc_expr[k], time, x[0], x[1], P->dim == 3 ? x[2] : 0.0, 0.0);
    return v;
}

static void gather(const orc_problem* P, const int64_t* nodes, const double* coords, int mode,
                   double time, const double* U, double* Ue) {
    for (int a = 0; a < P->npe; ++a) {
        const int32_t r = P->topo_to_alg[nodes[a]];
        Ue[a] = 0.0;
        if (r >= 0) { if (mode != 2 && U) Ue[a] = U[r]; }
        else if (mode != 0) Ue[a] = bc_value(P, P->ctag[nodes[a]], time, &coords[a * P->dim]);
    }
}

static void element_coords(const orc_problem* P, const int64_t* nodes, double* c) {
    for (int a = 0; a < P->npe; ++a)
        for (int i = 0; i < P->dim; ++i) c[a * P->dim + i] = P->xyz[nodes[a] * P->dim + i];
}

static int64_t elem_at(const orc_problem* P, int64_t k) { return P->elem_order ? P->elem_order[k] : k; }

int orc_assemble_matrix(const orc_problem* P, double time, int32_t n_forms, const int32_t* kinds,
                        const double* coefs, const double* U, const double* Udot,
                        const int64_t* map, int64_t nnz, double* val) {
    double xi[3 * 125], w[125], coords[3 * MAXNPE], Ke[MAXNPE * MAXNPE], Uv[MAXNPE], Ur[MAXNPE];
    const int nq = quad_points(P, xi, w), n = P->npe;
    for (int64_t i = 0; i < nnz; ++i) val[i] = 0.0;
    for (int64_t k = 0; k < P->n_elem; ++k) {
        const int64_t e = elem_at(P, k);
        const int64_t* nodes = P->ien + e * n;
        memset(Ke, 0, sizeof(Ke));
        element_coords(P, nodes, coords);
        gather(P, nodes, coords, 1, time, U, Uv);
        gather(P, nodes, coords, 0, time, Udot, Ur);
        for (int q = 0; q < nq; ++q) {
            qp_t Q;
            qp_evaluate(P, coords, &xi[P->dim * q], w[q], &Q);
            qp_interpolate(P, Uv, Ur, &Q);
            qp_models(P, e, &Q);
            for (int f = 0; f < n_forms; ++f) form_matrix(P, kinds[f], coefs[f], &Q, Ke);
        }
        const int64_t* m = map + e * n * n;
        for (int a = 0; a < n; ++a)
            for (int b = 0; b < n; ++b)
                if (m[a * n + b] >= 0) val[m[a * n + b]] += Ke[a * n + b];
    }
    return 0;
}

int orc_assemble_diagonal(const orc_problem* P, double time, int32_t n_forms, const int32_t* kinds,
                          const double* coefs, const double* U, double* diag) {
    double xi[3 * 125], w[125], coords[3 * MAXNPE], Ke[MAXNPE * MAXNPE], Uv[MAXNPE], Ur[MAXNPE];
    const int nq = quad_points(P, xi, w), n = P->npe;
    for (int64_t i = 0; i < P->n_free; ++i) diag[i] = 0.0;
    for (int64_t k = 0; k < P->n_elem; ++k) {
        const int64_t e = elem_at(P, k);
        const int64_t* nodes = P->ien + e * n;
        memset(Ke, 0, sizeof(Ke));
        element_coords(P, nodes, coords);
        gather(P, nodes, coords, 1, time, U, Uv);
        gather(P, nodes, coords, 0, time, NULL, Ur);
        for (int q = 0; q < nq; ++q) {
            qp_t Q;
            qp_evaluate(P, coords, &xi[P->dim * q], w[q], &Q);
            qp_interpolate(P, Uv, Ur, &Q);
            qp_models(P, e, &Q);
            for (int f = 0; f < n_forms; ++f) form_matrix(P, kinds[f], coefs[f], &Q, Ke);
        }
        for (int a = 0; a < n; ++a) {
            const int32_t r = P->topo_to_alg[nodes[a]];
            if (r >= 0) diag[r] += Ke[a * n + a];
        }
    }
    return 0;
}

int orc_assemble_vector(const orc_problem* P, double time, int32_t n_forms, const int32_t* kinds,
                        const double* coefs, const double* U, const double* Udot,
                        const double* operand, int32_t selector, double* F) {
    double xi[3 * 125], w[125], coords[3 * MAXNPE], Fe[MAXNPE], Uv[MAXNPE], Ur[MAXNPE], Uo[MAXNPE];
    const int nq = quad_points(P, xi, w), n = P->npe;
    for (int64_t i = 0; i < P->n_free; ++i) F[i] = 0.0;
    for (int64_t k = 0; k < P->n_elem; ++k) {
        const int64_t e = elem_at(P, k);
        const int64_t* nodes = P->ien + e * n;
        for (int a = 0; a < n; ++a) Fe[a] = 0.0;
        element_coords(P, nodes, coords);
        gather(P, nodes, coords, 1, time, U, Uv);
        gather(P, nodes, coords, 0, time, Udot, Ur);
        gather(P, nodes, coords, 0, time, operand, Uo);
        const double* sel = selector == ORC_SEL_VALUE ? Uv : (selector == ORC_SEL_RATE ? Ur : Uo);
        for (int q = 0; q < nq; ++q) {
            qp_t Q;
            qp_evaluate(P, coords, &xi[P->dim * q], w[q], &Q);
            qp_interpolate(P, Uv, Ur, &Q);
            qp_models(P, e, &Q);
            for (int f = 0; f < n_forms; ++f) form_vector(P, kinds[f], coefs[f], &Q, sel, Fe);
        }
        for (int a = 0; a < n; ++a) {
            const int32_t r = P->topo_to_alg[nodes[a]];
            if (r >= 0) F[r] += Fe[a];
        }
    }
    return 0;
}

int orc_assemble_load(const orc_problem* P, double time, const void* src_expr,
                      const double* nodal, double* F) {
    double xi[3 * 125], w[125], coords[3 * MAXNPE], Fe[MAXNPE], De[MAXNPE];
    const int nq = quad_points(P, xi, w), n = P->npe;
    for (int64_t i = 0; i < P->n_free; ++i) F[i] = 0.0;
    for (int64_t k = 0; k < P->n_elem; ++k) {
        const int64_t e = elem_at(P, k);
        const int64_t* nodes = P->ien + e * n;
        for (int a = 0; a < n; ++a) Fe[a] = 0.0;
        element_coords(P, nodes, coords);
        if (nodal) for (int a = 0; a < n; ++a) De[a] = nodal[nodes[a]];
        for (int q = 0; q < nq; ++q) {
            qp_t Q;
            qp_evaluate(P, coords, &xi[P->dim * q], w[q], &Q);
            double val = 0.0;
            if (nodal) {
                for (int b = 0; b < n; ++b) val += Q.N[b] * De[b];
            } else if (src_expr) {
                val = orc_expr_eval(src_expr, time, Q.x[0], Q.x[1], P->dim == 3 ? Q.x[2] : 0.0, 0.0);
            }
            for (int a = 0; a < n; ++a) Fe[a] += val * Q.N[a] * Q.measure * Q.w;
        }
        for (int a = 0; a < n; ++a) {
            const int32_t r = P->topo_to_alg[nodes[a]];
            if (r >= 0) F[r] += Fe[a];
        }
    }
    return 0;
}

int orc_apply_essential(const orc_problem* P, double time, int32_t n_forms, const int32_t* kinds,
                        const double* coefs, double* F) {
    double xi[3 * 125], w[125], coords[3 * MAXNPE], Fe[MAXNPE], Ge[MAXNPE];
    const int nq = quad_points(P, xi, w), n = P->npe;
    for (int64_t k = 0; k < P->n_elem; ++k) {
        const int64_t e = elem_at(P, k);
        const int64_t* nodes = P->ien + e * n;
        for (int a = 0; a < n; ++a) Fe[a] = 0.0;
        element_coords(P, nodes, coords);
        gather(P, nodes, coords, 2, time, NULL, Ge);
        for (int q = 0; q < nq; ++q) {
            qp_t Q;
            qp_evaluate(P, coords, &xi[P->dim * q], w[q], &Q);
            Q.T = 0.0; Q.Trate = 0.0; Q.gradT[0] = Q.gradT[1] = Q.gradT[2] = 0.0;
            qp_models(P, e, &Q);
            for (int f = 0; f < n_forms; ++f) form_vector(P, kinds[f], coefs[f], &Q, Ge, Fe);
        }
        for (int a = 0; a < n; ++a) {
            const int32_t r = P->topo_to_alg[nodes[a]];
            if (r >= 0) F[r] += -1.0 * Fe[a];
        }
    }
    return 0;
}

/* internal helpers for the drivers in oracle_solve.c */
void orc__gather(const orc_problem* P, int64_t e, int mode, double time, const double* U, double* Ue) {
    double coords[3 * MAXNPE];
    const int64_t* nodes = P->ien + e * P->npe;
    element_coords(P, nodes, coords);
    gather(P, nodes, coords, mode, time, U, Ue);
}

void orc__element_matrix(const orc_problem* P, int64_t e, double time, int32_t n_forms,
                         const int32_t* kinds, const double* coefs, const double* U, double* Ke) {
    double xi[3 * 125], w[125], coords[3 * MAXNPE], Uv[MAXNPE], Ur[MAXNPE];
    const int nq = quad_points(P, xi, w), n = P->npe;
    const int64_t* nodes = P->ien + e * n;
    memset(Ke, 0, sizeof(double) * (size_t)(n * n));
    element_coords(P, nodes, coords);
    gather(P, nodes, coords, 1, time, U, Uv);
    gather(P, nodes, coords, 0, time, NULL, Ur);
    for (int q = 0; q < nq; ++q) {
        qp_t Q;
        qp_evaluate(P, coords, &xi[P->dim * q], w[q], &Q);
        qp_interpolate(P, Uv, Ur, &Q);
        qp_models(P, e, &Q);
        for (int f = 0; f < n_forms; ++f) form_matrix(P, kinds[f], coefs[f], &Q, Ke);
    }
}

/* ------------------------------------------------------------------ */
/* weak boundary reaction (QuantityEvaluator<CPU>::evaluateBoundaryReaction,
 * fem/quantity/backend/cpu/QuantityEvaluator.tpp:95-199) with the heat
 * problem's pairs (HeatProblem.tpp:385-410): source on the value, -1 x
 * stiffness on the value, -1 x mass on the rate (transient only)          */
/* ------------------------------------------------------------------ */
/* face quadrature (GaussQuadrature1D / GaussQuadratureQuad, x fastest),
 * mapFaceToElement and getFaceTopology (LagrangeQuad.tpp:95-181,
 * LagrangeHex.tpp:120-267) */
static int face_points(const orc_problem* P, int f, double* xi_elem, double* w, double* nref) {
    double x[MAXQ1D], ww[MAXQ1D];
    const int d = P->dim, nq = P->nq[0];
    gauss1d(nq, x, ww);
    for (int i = 0; i < d; ++i) nref[i] = 0.0;
    nref[f / 2] = (f % 2) ? 1.0 : -1.0;
    const double side = (f % 2) ? 1.0 : -1.0;
    int q = 0;
    if (d == 2) {
        for (int i = 0; i < nq; ++i, ++q) {
            const double s = x[i];
            xi_elem[2 * q] = f < 2 ? side : s;
            xi_elem[2 * q + 1] = f < 2 ? s : side;
            w[q] = ww[i];
        }
        return q;
    }
    for (int j = 0; j < nq; ++j)
        for (int i = 0; i < nq; ++i, ++q) {
            const double s = x[i], t = x[j];
            double* e = xi_elem + 3 * q;
            if (f < 2) { e[0] = side; e[1] = s; e[2] = t; }
            else if (f < 4) { e[0] = s; e[1] = side; e[2] = t; }
            else { e[0] = s; e[1] = t; e[2] = side; }
            w[q] = ww[i] * ww[j];
        }
    return q;
}

/* JacobianTransform::computeBoundaryNormal for ParametricDim == SpatialDim
 * (JacobianTransform.tpp:110-123) with computeMatrixCofactor (:209-233) */
static void boundary_normal(int d, const double* A, const double* nref, double* n) {
    double c[9];
    if (d == 2) {
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
    for (int i = 0; i < d; ++i) {
        n[i] = 0.0;
        for (int j = 0; j < d; ++j) n[i] += c[i * d + j] * nref[j];
    }
}

int orc_boundary_reaction(const orc_problem* P, double time, const void* src_expr, const double* U,
                          const double* Udot, const int32_t* rng, int32_t n_tags, const int32_t* tags,
                          double* result, double* measure) {
    double xi[3 * 125], w[125], coords[3 * MAXNPE], Oe[MAXNPE], Uv[MAXNPE], Ur[MAXNPE];
    double fxi[3 * 25], fw[25], nref[3];
    int64_t fids[MAXNPE];
    int claimed[MAXNPE];
    const int nq = quad_points(P, xi, w), n = P->npe, d = P->dim, nf = 2 * d;
    for (int t = 0; t < n_tags; ++t) { result[t] = 0.0; measure[t] = 0.0; }
    for (int64_t e = 0; e < P->n_elem; ++e) {
        const int32_t* rt = rng + e * nf;
        int touches = 0;
        for (int f = 0; f < nf; ++f)
            for (int t = 0; t < n_tags; ++t) touches |= rt[f] == tags[t];
        if (!touches) continue;
        const int64_t* nodes = P->ien + e * n;
        for (int a = 0; a < n; ++a) { Oe[a] = 0.0; claimed[a] = -1; }
        element_coords(P, nodes, coords);
        gather(P, nodes, coords, 1, time, U, Uv);
        gather(P, nodes, coords, 0, time, Udot, Ur);
        for (int q = 0; q < nq; ++q) {
            qp_t Q;
            qp_evaluate(P, coords, &xi[d * q], w[q], &Q);
            qp_interpolate(P, Uv, Ur, &Q);
            qp_models(P, e, &Q);
            /* SourceForm (SourceForm.hpp:22-32) */
            const double fval = src_expr ? orc_expr_eval(src_expr, time, Q.x[0], Q.x[1], d == 3 ? Q.x[2] : 0.0, 0.0) : 0.0;
            for (int a = 0; a < n; ++a) Oe[a] += fval * Q.N[a] * Q.measure * Q.w;
            form_vector(P, ORC_FORM_DIFFUSION | ORC_SCALED, -1.0, &Q, Uv, Oe);
            if (Udot) form_vector(P, ORC_FORM_MASS | ORC_SCALED, -1.0, &Q, Ur, Oe);
        }
        for (int f = 0; f < nf; ++f) {
            int t = -1;
            for (int k = 0; k < n_tags; ++k) if (rt[f] == tags[k]) t = k;
            if (t < 0) continue;
            const int fq = face_points(P, f, fxi, fw, nref);
            for (int q = 0; q < fq; ++q) {
                qp_t Q;
                qp_evaluate(P, coords, &fxi[d * q], fw[q], &Q);
                double nv[3], mag = 0.0;
                boundary_normal(d, Q.J, nref, nv);
                for (int s = 0; s < d; ++s) mag += nv[s] * nv[s];
                measure[t] += sqrt(mag) * fw[q];
            }
            const int nfn = face_nodes(d, P->p, f, fids);
            for (int i = 0; i < nfn; ++i)
                if (claimed[fids[i]] != rt[f]) {
                    result[t] += Oe[fids[i]];
                    claimed[fids[i]] = rt[f];
                }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* strong (pointwise) boundary heat flux: QuantityEvaluator<CPU>::
 * evaluateBoundaryRegistry (fem/quantity/backend/cpu/QuantityEvaluator.tpp:
 * 14-93) with HeatFluxIntegrand (equation/heateq/quantity/
 * HeatFluxIntegrand.hpp:19-38) at the boundary points of
 * EvalQuadraturePointBoundary::evaluate (equation/heateq/evaluator/
 * EvalQuadraturePointBoundary.hpp:73-101): per face point the full-gathered
 * T gives gradT and k(T), and the term -(K gradT).n_cof * w; the measure
 * terms as in the weak reaction.  Integral reduction (the Average division
 * by the measure is the caller's).                                        */
/* ------------------------------------------------------------------ */
int orc_boundary_flux(const orc_problem* P, double time, const double* U, const int32_t* rng, int32_t n_tags,
                      const int32_t* tags, double* result, double* measure) {
    double coords[3 * MAXNPE], Uv[MAXNPE], zero[MAXNPE];
    double fxi[3 * 25], fw[25], nref[3];
    const int n = P->npe, d = P->dim, nf = 2 * d;
    for (int a = 0; a < n; ++a) zero[a] = 0.0;
    for (int t = 0; t < n_tags; ++t) { result[t] = 0.0; measure[t] = 0.0; }
    for (int64_t e = 0; e < P->n_elem; ++e) {
        const int32_t* rt = rng + e * nf;
        int touches = 0;
        for (int f = 0; f < nf; ++f)
            for (int t = 0; t < n_tags; ++t) touches |= rt[f] == tags[t];
        if (!touches) continue;
        const int64_t* nodes = P->ien + e * n;
        element_coords(P, nodes, coords);
        gather(P, nodes, coords, 1, time, U, Uv);
        for (int f = 0; f < nf; ++f) {
            int t = -1;
            for (int k = 0; k < n_tags; ++k) if (rt[f] == tags[k]) t = k;
            if (t < 0) continue;
            const int fq = face_points(P, f, fxi, fw, nref);
            for (int q = 0; q < fq; ++q) {
                qp_t Q;
                qp_evaluate(P, coords, &fxi[d * q], fw[q], &Q);
                qp_interpolate(P, Uv, zero, &Q);
                qp_models(P, e, &Q);
                double nv[3], mag = 0.0, flux = 0.0;
                boundary_normal(d, Q.J, nref, nv);
                for (int i = 0; i < d; ++i) {
                    double kg = 0.0;
                    for (int j = 0; j < d; ++j) kg += Q.K[i * d + j] * Q.gradT[j];
                    flux -= kg * nv[i];
                }
                result[t] += flux * fw[q];
                for (int s = 0; s < d; ++s) mag += nv[s] * nv[s];
                measure[t] += sqrt(mag) * fw[q];
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* natural (flux) boundary conditions: BoundaryApplicator<CPU>::
 * applyNaturalBCs (fem/boundary/backend/cpu/BoundaryApplicator.tpp:66-140)
 * with FluxBoundaryForm (equation/heateq/form/FluxBoundaryForm.hpp:24-36)
 * or NodalFluxForm (NodalFluxForm.hpp:28-51) at the points of
 * EvalQuadraturePointBoundary::evaluate (EvalQuadraturePointBoundary.hpp:
 * 73-101; Nface = the volume shape values at the face nodes).  Elements in
 * ascending order, faces ascending; a face whose tag has entries gets
 * Fe (zeroed), every point adds each entry of its tag in order,
 *   Fe[a] += (q_sD Nface_a) n_sD w  (sD ascending),
 * then Fe is scattered into F at the free face nodes (scatterElementVector,
 * ElementMap.tpp:77-92, coefficient 1).  Entry k: nodal[k] != NULL gives
 * nodal flux vectors [n_nodes * dim] (q = sum_b Nface_b q_b from 0), else
 * the dim expressions exprs[k * dim + sD] of (t, x, y, z).                */
/* ------------------------------------------------------------------ */
int orc_apply_natural(const orc_problem* P, double time, const int32_t* rng, int32_t n_entries,
                      const int32_t* entry_tags, const void* const* exprs, const double* const* nodal, double* F) {
    double coords[3 * MAXNPE], Fe[MAXNPE], fxi[3 * 25], fw[25], nref[3];
    int64_t fids[MAXNPE];
    const int n = P->npe, d = P->dim, nf = 2 * d;
    for (int64_t e = 0; e < P->n_elem; ++e) {
        const int64_t* nodes = P->ien + e * n;
        const int32_t* rt = rng + e * nf;
        int any = 0;
        for (int f = 0; f < nf; ++f)
            for (int k = 0; k < n_entries; ++k) any |= rt[f] >= 0 && rt[f] == entry_tags[k];
        if (!any) continue;
        element_coords(P, nodes, coords);
        for (int f = 0; f < nf; ++f) {
            int has = 0;
            for (int k = 0; k < n_entries; ++k) has |= rt[f] >= 0 && rt[f] == entry_tags[k];
            if (!has) continue;
            const int npf = face_nodes(d, P->p, f, fids);
            for (int a = 0; a < npf; ++a) Fe[a] = 0.0;
            const int fq = face_points(P, f, fxi, fw, nref);
            for (int q = 0; q < fq; ++q) {
                qp_t Q;
                qp_evaluate(P, coords, &fxi[d * q], fw[q], &Q);
                double nv[3];
                boundary_normal(d, Q.J, nref, nv);
                for (int k = 0; k < n_entries; ++k) {
                    if (entry_tags[k] != rt[f]) continue;
                    double val[3] = {0.0, 0.0, 0.0};
                    if (nodal && nodal[k]) {
                        for (int b = 0; b < npf; ++b)
                            for (int s = 0; s < d; ++s) val[s] += Q.N[fids[b]] * nodal[k][nodes[fids[b]] * d + s];
                    } else {
                        for (int s = 0; s < d; ++s)
                            val[s] = orc_expr_eval(exprs[k * d + s], time, Q.x[0], Q.x[1], d == 3 ? Q.x[2] : 0.0, 0.0);
                    }
                    for (int a = 0; a < npf; ++a)
                        for (int s = 0; s < d; ++s) Fe[a] += ((val[s] * Q.N[fids[a]]) * nv[s]) * fw[q];
                }
            }
            for (int a = 0; a < npf; ++a) {
                const int32_t r = P->topo_to_alg[nodes[fids[a]]];
                if (r >= 0) F[r] += 1.0 * Fe[a];
            }
        }
    }
    return 0;
}
