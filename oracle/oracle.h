/*
 * oracle.h -- CPU FP64 oracle for the Residuum heat-equation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a serial, plain-C restatement
 * of the reference algorithms (elicase4/residuum-solver).  Only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark may load
 * it, and only as the checker: no product code path calls into it.
 *
 * Every function names the reference file:line it restates.  Two orderings
 * are provided where floating-point summation order matters:
 *   ORC_ORDER_REF  -- the reference's own order (ascending element loop,
 *                     left-to-right sequential dot products);
 *   ORC_ORDER_GPU  -- the order the HIP kernels use (colour-major element
 *                     loop, fixed 256-wide block-tree reductions), so the
 *                     GPU results can be checked bit for bit.
 *
 * Parity pins (see tests/):
 *   mesh, constraints, CSR pattern, elem->nnz map: bit-exact vs golden
 *     vectors produced by oracle/_ref (the reference's own headers) and vs
 *     brute-force definitions;
 *   element math / assembly / lift: golden vectors from oracle/_ref and the
 *     closed-form values of the reference's fem-tier tests;
 *   CG / GMRES / Jacobi: golden iterates from oracle/_ref, known 2x2
 *     solutions, exact bilinear solutions;
 *   Newton: Kirchhoff closed form; BE: eigenmode decay, linear-field
 *     preservation; adjoint: central finite differences and the discrete
 *     adjoint identity.
 */
#ifndef RESIDUUM_ORACLE_H
#define RESIDUUM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_ORDER_REF = 0, ORC_ORDER_GPU = 1 };

/* form kinds, applied in list order (fem/form/FormRegistry.hpp:202-216).
 * OR-ing ORC_SCALED wraps the form in a ScaledForm (fem/form/ScaledForm.hpp:
 * 14-57) with the matching coefficient; unscaled forms ignore it. */
enum { ORC_FORM_DIFFUSION = 1, ORC_FORM_MASS = 2, ORC_FORM_TDIFFUSION = 3, ORC_FORM_TMASS = 4 };
#define ORC_SCALED 0x100

/* state buffer selector for vector assembly (HeatStateVector.hpp:28-30) */
enum { ORC_SEL_VALUE = 0, ORC_SEL_RATE = 1, ORC_SEL_OPERAND = 2 };

/* ------------------------------------------------------------------ */
/* expressions: x, y, z, t, T; + - * / ^; sin cos tan exp log sqrt abs */
/* pow min max; pi e.  Stand-in for the reference's exprtk call sites   */
/* (src/utils/expression/ScalarExpression.cpp:3-30).                    */
/* ------------------------------------------------------------------ */
void*  orc_expr_compile(const char* text);
void   orc_expr_free(void* e);
double orc_expr_eval(const void* e, double t, double x, double y, double z, double T);

/* ------------------------------------------------------------------ */
/* problem descriptor                                                   */
/* ------------------------------------------------------------------ */
#define ORC_MAX_BC 16
typedef struct orc_problem {
    int32_t dim;                /* 2 = quad, 3 = hex */
    int32_t p[3];               /* basis order per direction */
    int32_t nq[3];              /* Gauss points per direction */
    int32_t npe;                /* nodes per element */
    int64_t n_nodes;
    int64_t n_elem;
    int64_t n_free;
    const double*  xyz;         /* [n_nodes*dim] */
    const int64_t* ien;         /* [n_elem*npe] */
    const int32_t* topo_to_alg; /* [n_nodes], -1 = constrained */
    const int32_t* ctag;        /* [n_nodes], constraint tag or -1 */
    const int64_t* elem_order;  /* element visiting order, NULL = ascending */
    /* conductivity K = kappa_e * k(T) * shape, dK/dT = kappa_e * k'(T) * shape */
    const double* kappa_e;      /* [n_elem] or NULL */
    const void*   k_expr;       /* k(T) or NULL -> k_scalar */
    const void*   dk_expr;      /* k'(T) or NULL -> 0 */
    double  k_scalar;
    int32_t anisotropic;
    double  k_tensor[9];
    /* mass model */
    double  rho;
    double  cp;
    const void* cp_expr;        /* cp(T) or NULL -> cp */
    const void* dcp_expr;       /* cp'(T) or NULL -> 0 */
    /* Dirichlet registry: tag -> value expression g(t, x, y, z) */
    int32_t n_bc;
    int32_t bc_tag[ORC_MAX_BC];
    const void* bc_expr[ORC_MAX_BC];
} orc_problem;

/* ------------------------------------------------------------------ */
/* mesh (src/mesh/generator/BlockMesh2D.cpp, BlockMesh3D.cpp)           */
/* ------------------------------------------------------------------ */
void orc_block_mesh_2d(int64_t nx, int64_t ny, double x0, double x1, double y0, double y1,
                       int64_t px, int64_t py, double* xyz, int64_t* ien, int32_t* rng);
void orc_block_mesh_3d(int64_t nx, int64_t ny, int64_t nz, double x0, double x1, double y0,
                       double y1, double z0, double z1, int64_t px, int64_t py, int64_t pz,
                       double* xyz, int64_t* ien, int32_t* rng);

/* constraints + interleaved numbering (topology/TopologicalDOF.tpp:111-190).
 * returns the number of free DOFs */
int64_t orc_build_constraints(int32_t dim, const int32_t* p, int64_t n_nodes, int64_t n_elem,
                              const int64_t* ien, const int32_t* rng, int32_t n_tags,
                              const int32_t* tags, int32_t* topo_to_alg, int32_t* alg_to_topo,
                              int32_t* ctag);

/* CSR pattern (fem/assembly/backend/cpu/Assembler.tpp:7-76).  First call
 * with col_idx == NULL to get nnz and row_ptr, then again with col_idx. */
int64_t orc_create_matrix(int64_t n_elem, int32_t npe, const int64_t* ien,
                          const int32_t* topo_to_alg, int64_t n_free, int64_t* row_ptr,
                          int64_t* col_idx);

/* element -> nnz map, -1 where row or column is constrained
 * (ElementMap.tpp:115-184 + CSRMatrix.hpp:49-73 getDataIndex) */
void orc_elem_nnz_map(int64_t n_elem, int32_t npe, const int64_t* ien, const int32_t* topo_to_alg,
                      const int64_t* row_ptr, const int64_t* col_idx, int64_t* map);

/* first-fit greedy element colouring in ascending element order (no two
 * elements of one colour share a node), then the colour-major element
 * order the HIP assembly kernels visit.  On block meshes this reproduces
 * the parity colouring.  Returns the number of colours. */
int32_t orc_colour_order(int64_t n_elem, int32_t npe, const int64_t* ien, int64_t n_nodes,
                         int32_t* colour, int64_t* order);

/* ------------------------------------------------------------------ */
/* assembly                                                             */
/* ------------------------------------------------------------------ */
/* Assembler<CPU>::assembleMatrix (Assembler.tpp:86-140): val zeroed then
 * accumulated.  U = value source (free vector, Full gather incl. BCs),
 * Udot = rate source or NULL. */
int orc_assemble_matrix(const orc_problem* P, double time, int32_t n_forms, const int32_t* kinds,
                        const double* coefs, const double* U, const double* Udot,
                        const int64_t* map, int64_t nnz, double* val);

/* Assembler<CPU>::assembleVector (Assembler.tpp:199-255) with one
 * FormsWithOperand pair.  operand may be NULL. */
int orc_assemble_vector(const orc_problem* P, double time, int32_t n_forms, const int32_t* kinds,
                        const double* coefs, const double* U, const double* Udot,
                        const double* operand, int32_t selector, double* F);

/* load vector: SourceForm (SourceForm.hpp:22-32) when src_expr != NULL,
 * NodalSourceForm (NodalSourceForm.hpp:62-86) when nodal != NULL */
int orc_assemble_load(const orc_problem* P, double time, const void* src_expr,
                      const double* nodal, double* F);

/* BoundaryApplicator<CPU>::applyEssentialBCs (BoundaryApplicator.tpp:7-64) */
int orc_apply_essential(const orc_problem* P, double time, int32_t n_forms, const int32_t* kinds,
                        const double* coefs, double* F);

/* Assembler<CPU>::assembleDiagonal (Assembler.tpp:142-197) */
int orc_assemble_diagonal(const orc_problem* P, double time, int32_t n_forms, const int32_t* kinds,
                          const double* coefs, const double* U, double* diag);

/* QuantityEvaluator<CPU>::evaluateBoundaryReaction
 * (fem/quantity/backend/cpu/QuantityEvaluator.tpp:95-199) with the weak
 * heat-flux pairs of HeatProblem.tpp:385-410: per registered tag t, the sum
 * of the element vectors (source - K(U) U [- M Udot when Udot != NULL]) at
 * the nodes of the element's faces tagged t (each node once per element and
 * tag), in ascending element order, and the faces' measure.  U is the free
 * vector (Dirichlet values from the BCs at `time`); rng [n_elem * 2 dim]. */
int orc_boundary_reaction(const orc_problem* P, double time, const void* src_expr, const double* U,
                          const double* Udot, const int32_t* rng, int32_t n_tags, const int32_t* tags,
                          double* result, double* measure);
/* QuantityEvaluator<CPU>::evaluateBoundaryRegistry (QuantityEvaluator.tpp:
 * 14-93) with HeatFluxIntegrand (HeatFluxIntegrand.hpp:19-38), Integral
 * reduction: per tag the sum over its faces' points of -(K gradT).n w, n the
 * unnormalised cofactor normal, in (element, face, point) order, and the
 * faces' measure.  T is the full gather of U at `time`. */
int orc_boundary_flux(const orc_problem* P, double time, const double* U, const int32_t* rng, int32_t n_tags,
                      const int32_t* tags, double* result, double* measure);

/* ------------------------------------------------------------------ */
/* linear algebra                                                       */
/* ------------------------------------------------------------------ */
void   orc_spmv(int64_t n, const int64_t* row_ptr, const int64_t* col_idx, const double* val,
                const double* x, double* y);                 /* MatrixOps.tpp:5-23 */
double orc_dot(int64_t n, const double* x, const double* y, int32_t order); /* VectorOps.tpp:7-18 */
void   orc_inv_diag(int64_t n, const int64_t* row_ptr, const int64_t* col_idx, const double* val,
                    double* inv_diag);                       /* Jacobi.hpp:23-36 */

typedef struct orc_report {
    int32_t converged;
    int64_t iterations;
    double  initial_residual;
    double  final_residual;
    double  final_residual_rel;
} orc_report;

/* cg::Solver::solve (linalg/solver/iterative/cg/Solver.tpp:15-119).
 * precond: 0 identity, 1 Jacobi.  hist (nullable, >= max_iters entries)
 * receives ||r_k||. */
int orc_cg(int64_t n, const int64_t* row_ptr, const int64_t* col_idx, const double* val,
           const double* b, double* x, double tol, int64_t max_iters, int32_t precond,
           int32_t order, orc_report* rep, double* hist);

/* cg::Solver::solve over the matrix-free FEMOperator
 * (linalg/operator/FEMOperator.hpp:38-56): apply = assembleVector with the
 * operand selector, Jacobi from assembleDiagonal (Assembler.tpp:142-197). */
int orc_cg_fem(const orc_problem* P, double time, int32_t n_forms, const int32_t* kinds, const double* coefs,
               const double* U, const double* Udot, const double* b, double* x, double tol, int64_t max_iters,
               int32_t precond, int32_t order, orc_report* rep, double* hist);

/* gmres::Solver::solve (linalg/solver/iterative/gmres/Solver.tpp:5-170) */
int orc_gmres(int64_t n, const int64_t* row_ptr, const int64_t* col_idx, const double* val,
              const double* b, double* x, double tol, int64_t max_iters, int64_t krylov,
              int32_t precond, int32_t order, orc_report* rep);

/* ------------------------------------------------------------------ */
/* drivers                                                              */
/* ------------------------------------------------------------------ */
/* Linear Backward Euler (BackwardEulerStage.hpp:221-258,
 * BackwardEuler.hpp:132-172, Transient.hpp:248-262) with CSR + CG.
 * U holds the initial free vector and receives the final one.  U_hist
 * (nullable) receives U after each step ([n_steps_max * n_free]).
 * iters (nullable) receives the CG iteration count per step.
 * Returns the number of steps taken, or -1 if a solve failed. */
int64_t orc_be_run(const orc_problem* P, double t0, double tf, double dt, const void* src_expr,
                   const int64_t* row_ptr, const int64_t* col_idx, const int64_t* map, int64_t nnz,
                   double cg_tol, int64_t cg_max_iters, int32_t precond, int32_t order,
                   double* U, double* U_hist, int64_t n_steps_max, int64_t* iters);

/* Steady linear solve (SteadyStage.hpp:60-84, Steady.hpp:19-27):
 * assemble K (diffusion), load, lift, CG or GMRES (solver 0 / 1). */
int orc_steady_linear(const orc_problem* P, const void* src_expr, const double* nodal_src,
                      const int64_t* row_ptr, const int64_t* col_idx, const int64_t* map,
                      int64_t nnz, int32_t solver, double tol, int64_t max_iters, int64_t krylov,
                      int32_t precond, int32_t order, double* U, orc_report* rep);

/* Steady Newton (Newton.hpp:32-83, SteadyStage.hpp:60-116) with GMRES or
 * CG inner solves; returns Newton iterations (<0 on failure). res_hist
 * (nullable, max_newton entries) receives ||R||. */
int64_t orc_newton_steady(const orc_problem* P, const void* src_expr, const int64_t* row_ptr,
                          const int64_t* col_idx, const int64_t* map, int64_t nnz,
                          double abs_tol, double rel_tol, int64_t max_newton, int32_t lin_solver,
                          double lin_tol, int64_t lin_max_iters, int64_t krylov, int32_t precond,
                          int32_t order, double* U, double* res_hist, int64_t* lin_iters);

/* Adjoint gradient of 0.5*||U_N[obs] - T_obs||^2 with respect to the
 * per-element log-conductivity theta (kappa_e = exp(theta_e)) through
 * n_steps linear BE steps of fixed dt from U0 (no reference counterpart).
 * P->kappa_e must equal exp(theta).  Returns the loss; grad receives
 * [n_elem]; U_final (nullable) receives U_N. */
double orc_adjoint_grad(const orc_problem* P, double t0, double dt, int64_t n_steps,
                        const int64_t* row_ptr, const int64_t* col_idx, const int64_t* map,
                        int64_t nnz, double cg_tol, int64_t cg_max_iters, int32_t order,
                        const double* U0, int64_t n_obs, const int64_t* obs, const double* t_obs,
                        double* grad, double* U_final);

/* natural (flux) boundary conditions, BoundaryApplicator.tpp:66-140 with
 * FluxBoundaryForm / NodalFluxForm: F += each tagged face's vector; entry k
 * applies to faces tagged entry_tags[k], its flux from nodal[k] ([n_nodes *
 * dim]) when nodal and nodal[k] are non-NULL, else from the dim compiled
 * expressions exprs[k * dim + sD] */
int orc_apply_natural(const orc_problem* P, double time, const int32_t* rng, int32_t n_entries,
                      const int32_t* entry_tags, const void* const* exprs, const double* const* nodal, double* F);

/* ------------------------------------------------------------------ */
/* seeded inputs                                                        */
/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11) -> N(0,1) via Box-Muller.
 * out[i] = normal sample for counter (i, stream) with key seed. */
void orc_normal_field(uint64_t seed, uint32_t stream, int64_t n, double* out);
/* the bare Philox4x32-10 block: out = philox(ctr, key), ten rounds */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_uniform_field(uint64_t seed, uint32_t stream, int64_t n, double* out);

#ifdef __cplusplus
}
#endif
#endif
