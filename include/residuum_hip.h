/*
 * residuum_hip.h -- C-ABI of libresiduum_hip.so, the MI355X (gfx950) backend
 * for Residuum's heat-equation hot path.
 *
 * Every entry point returns an int status: RSD_OK (0) or a negative error
 * code; rsd_last_error_string() describes the last failure on the calling
 * thread.  Sizes are int64_t; device pointers are plain pointers obtained
 * from rsd_malloc (or any hipMalloc'd memory on the context's device).  All
 * work is ordered on the context's HIP stream; functions that return host
 * values synchronise that stream.  A context is not re-entrant: use one per
 * host thread.  Results are deterministic for identical inputs.
 *
 * Each declaration names the reference interface it replaces
 * (elicase4/residuum-solver, path:line).  The reference has no GPU code and no
 * FFI; these functions are what its Backend-tag seam would bind
 * (linalg/types/backend/CPU.hpp:15-40 and the CUDA tag scaffold in
 * linalg/types/backend/CUDA.hpp:16-73).
 */
#ifndef RESIDUUM_HIP_H
#define RESIDUUM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* status codes                                                            */
/* ---------------------------------------------------------------------- */
#define RSD_OK 0
#define RSD_ERR_ARG (-1)        /* invalid argument or unsupported discretisation */
#define RSD_ERR_HIP (-2)        /* HIP runtime error */
#define RSD_ERR_INTERNAL (-3)
#define RSD_ERR_NOT_CONVERGED (-4)

/* copy kinds: linalg/types/backend/cuda/CopyKind.hpp:9-14 */
#define RSD_COPY_H2H 0
#define RSD_COPY_H2D 1
#define RSD_COPY_D2H 2
#define RSD_COPY_D2D 3

/* form kinds (fem/form/FormRegistry.hpp:202-216); OR RSD_FORM_SCALED to wrap
 * a form in fem/form/ScaledForm.hpp:14-57 with the matching coefficient */
#define RSD_FORM_DIFFUSION 1        /* equation/heateq/form/DiffusionForm.hpp:11-52 */
#define RSD_FORM_MASS 2             /* equation/heateq/form/MassForm.hpp:11-27 */
#define RSD_FORM_TDIFFUSION 3       /* equation/heateq/form/TangentDiffusionForm.hpp:11-56 */
#define RSD_FORM_TMASS 4            /* equation/heateq/form/TangentMassForm.hpp:11-27 */
#define RSD_FORM_SCALED 0x100

/* state buffer selectors: equation/heateq/evaluator/HeatStateVector.hpp:28-30 */
#define RSD_SEL_VALUE 0
#define RSD_SEL_RATE 1
#define RSD_SEL_OPERAND 2

/* preconditioners: linalg/solver/preconditioner/{Identity,Jacobi}.hpp */
#define RSD_PC_IDENTITY 0
#define RSD_PC_JACOBI 1

typedef struct rsd_ctx rsd_ctx;
typedef struct rsd_problem rsd_problem;
typedef struct rsd_matrix rsd_matrix;
typedef struct rsd_batch rsd_batch;
typedef struct rsd_expr rsd_expr;

/* ---------------------------------------------------------------------- */
/* context                                                                 */
/* ---------------------------------------------------------------------- */
/* one HIP stream + reduction scratch on `device`.  Problems, matrices,
 * batches and rsd_malloc buffers made on a context are destroyed before it:
 * their destroy calls synchronise on the context's stream. */
int rsd_ctx_create(int device, rsd_ctx** out);
int rsd_ctx_destroy(rsd_ctx* ctx);
int rsd_ctx_sync(rsd_ctx* ctx);
/* the context's stream, as a hipStream_t, for callers that overlap work */
int rsd_ctx_stream(rsd_ctx* ctx, void** stream);
/* GMRES orthogonalisation for the context's later GMRES solves
 * (rsd_gmres_solve_f64, rsd_gmres_solve_mixed's FP32 inner GMRES, Newton's
 * GMRES inner solves; no reference counterpart).  RSD_ORTHO_MGS (the default): the reference's modified
 * Gram-Schmidt (gmres/Solver.tpp:71-78), bitwise, one fold of block sums
 * per basis vector.  RSD_ORTHO_CGS2: classical Gram-Schmidt with one
 * reorthogonalisation -- h = Q^T w, w -= Q h, c = Q^T w, w -= Q c,
 * H[:, j] = h + c -- three passes over the basis and three folds per Arnoldi
 * step; iterates differ from the reference's by rounding. */
#define RSD_ORTHO_MGS 0
#define RSD_ORTHO_CGS2 1
int rsd_ctx_set_gmres_orthogonalization(rsd_ctx* ctx, int32_t mode);
const char* rsd_last_error_string(void);
const char* rsd_version(void);
/* build stamp (no reference counterpart): the first 16 hex digits of the
 * SHA-256 of this header as the library was compiled against it; a binding
 * compares it with the header it reads, so a stale library is refused */
const char* rsd_abi_hash(void);

/* ---------------------------------------------------------------------- */
/* Backend-tag primitives: CPU.hpp:15-40 alloc/copy/zero/set               */
/* (CUDA.hpp:30-64 cudaMalloc / cudaMemcpy / cudaMemset / cudaFree)        */
/* ---------------------------------------------------------------------- */
int rsd_malloc(rsd_ctx* ctx, size_t bytes, void** out);
/* *out = 1 when p is device memory (rsd_malloc / hipMalloc), 0 for host
 * memory: lets host code that reads a vector's data() (the reference's
 * fem/dof/DOFResidualNorms.hpp:14-53 behind its loggers and NewtonRunner)
 * tell a Hip vector from a CPU one */
int rsd_pointer_is_device(const void* p, int32_t* out);
int rsd_free(rsd_ctx* ctx, void* p);
int rsd_memcpy(rsd_ctx* ctx, void* dst, const void* src, size_t bytes, int kind);
int rsd_memset_zero(rsd_ctx* ctx, void* p, size_t bytes);
/* Vector::set (linalg/types/Vector.hpp:37-39): a fill kernel, which memset
 * cannot do for doubles (CUDA.hpp:66-71 leaves it commented out) */
int rsd_fill_f64(rsd_ctx* ctx, double* p, int64_t n, double value);
int rsd_fill_f32(rsd_ctx* ctx, float* p, int64_t n, float value);

/* ---------------------------------------------------------------------- */
/* BLAS-1: linalg/operations/backend/cpu/VectorOps.tpp:7-86                */
/* dot/norm write the scalar to d_out (device, nullable) and h_out (host,  */
/* nullable; synchronises).  Reductions use a fixed tree: 256-index       */
/* blocks, 64-lane trees per wavefront, wave sums left to right, block    */
/* partials gathered per lane and reduced the same way, so results are    */
/* bitwise reproducible run to run.                                        */
/* ---------------------------------------------------------------------- */
int rsd_dot_f64(rsd_ctx* ctx, int64_t n, const double* x, const double* y, double* d_out, double* h_out);
int rsd_norm2_f64(rsd_ctx* ctx, int64_t n, const double* x, double* d_out, double* h_out);
int rsd_axpy_f64(rsd_ctx* ctx, int64_t n, double alpha, const double* x, double* y);           /* y += a x */
int rsd_axpby_f64(rsd_ctx* ctx, int64_t n, double alpha, double beta, const double* x, double* y); /* y = b y + a x */
int rsd_scal_f64(rsd_ctx* ctx, int64_t n, double alpha, double* x);
int rsd_copy_f64(rsd_ctx* ctx, int64_t n, const double* x, double* y);
int rsd_mul_f64(rsd_ctx* ctx, int64_t n, const double* x, const double* y, double* z);        /* z = x .* y */
/* rsd_dot_f64 for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops); no counterpart in the reference.  The same
 * kernels and result bits as rsd_dot_f64, the block partials in a static
 * 512 KB device array instead of the context's partials area: no allocation
 * on any call.  d_out (device) receives x . y; nothing is synchronised.
 * n <= 2^24, else RSD_ERR_ARG.  The array is one per device, shared by every
 * call on that device: calls that overlap on one device (other streams or
 * contexts) race. */
int rsd_dot_f64_noalloc(rsd_ctx* ctx, int64_t n, const double* x, const double* y, double* d_out);
/* The same operations on FP32 vectors (Vector<float>, VectorOps.tpp:7-86):
 * the elementwise ones in float, each product and sum rounded once as in the
 * CPU loop (bitwise it); dot and norm2 form each product exactly in FP64 and
 * sum with the FP64 tree above, writing a double (the CPU loop accumulates
 * in float, so the two agree to float rounding). */
int rsd_dot_f32(rsd_ctx* ctx, int64_t n, const float* x, const float* y, double* d_out, double* h_out);
int rsd_norm2_f32(rsd_ctx* ctx, int64_t n, const float* x, double* d_out, double* h_out);
int rsd_axpy_f32(rsd_ctx* ctx, int64_t n, float alpha, const float* x, float* y);
int rsd_axpby_f32(rsd_ctx* ctx, int64_t n, float alpha, float beta, const float* x, float* y);
int rsd_scal_f32(rsd_ctx* ctx, int64_t n, float alpha, float* x);
int rsd_copy_f32(rsd_ctx* ctx, int64_t n, const float* x, float* y);
int rsd_mul_f32(rsd_ctx* ctx, int64_t n, const float* x, const float* y, float* z);

/* ---------------------------------------------------------------------- */
/* problem: mesh + topological DOFs + CSR pattern + element->nnz map +     */
/* colouring + material model, resident in HBM                             */
/* ---------------------------------------------------------------------- */
/* BlockMesh2D/3D::generate (src/mesh/generator/BlockMesh{2D,3D}.cpp),
 * TopologicalDOF::buildConstraints (topology/TopologicalDOF.tpp:111-190),
 * Assembler<CPU>::createMatrix (fem/assembly/backend/cpu/Assembler.tpp:7-76)
 * and the implicit element->nnz map of scatterElementMatrix
 * (ElementMap.tpp:115-184).  dim 2 (quad) or 3 (hex); n[3], p[3] element
 * counts and basis orders; ext[6] = x0 x1 y0 y1 z0 z1; nq Gauss points per
 * direction; Dirichlet tags with value expressions g(t, x, y, z). */
int rsd_problem_create_block(rsd_ctx* ctx, int32_t dim, const int64_t* n, const double* ext,
                             const int32_t* p, int32_t nq, int32_t n_bc, const int32_t* bc_tags,
                             const char* const* bc_exprs, rsd_problem** out);
/* BlockMesh2D/3D::generate on the host (no device work): xyz
 * [n_nodes*dim], ien [n_elem*npe], rng [n_elem*2*dim] with n_nodes =
 * prod(n_i*p+1), n_elem = prod(n_i); one basis order p[0] == p[1] (== p[2]) */
int rsd_block_mesh(int32_t dim, const int64_t* n, const double* ext, const int32_t* p, double* xyz, int64_t* ien,
                   int32_t* rng);
/* same from host mesh arrays (mesh/Data.hpp:13-38 layout: xyz [nn*dim],
 * ien [ne*npe], rng [ne*2*dim]) */
int rsd_problem_create(rsd_ctx* ctx, int32_t dim, const int32_t* p, int32_t nq, int64_t n_nodes,
                       int64_t n_elem, const double* xyz, const int64_t* ien, const int32_t* rng,
                       int32_t n_bc, const int32_t* bc_tags, const char* const* bc_exprs, rsd_problem** out);
/* Destroying a problem orphans the matrices made on it: every later call
 * with such a matrix fails with RSD_ERR_ARG, and rsd_matrix_destroy still
 * releases it. */
int rsd_problem_destroy(rsd_problem* pb);
/* the Gauss rule (nq points per direction, 1..5) of the element and face
 * integrals from now on; the mesh, DOF maps and pattern do not depend on it
 * (the reference passes the rule to each assembly call, Assembler.hpp:58-66) */
int rsd_problem_set_quadrature(rsd_ctx* ctx, rsd_problem* pb, int32_t nq);
/* out[8] = n_nodes, n_elem, npe, n_free, nnz, n_colours, sell_slots, max_row */
int rsd_problem_sizes(rsd_problem* pb, int64_t* out);
int rsd_problem_get_mesh(rsd_problem* pb, double* xyz, int64_t* ien, int32_t* rng);
/* matrix layout (no reference counterpart): out[4] = column mode (0 absolute
 * int32 columns, 1 int16 row-relative offsets, 2 offset patterns shared by
 * rows: a pattern id per row, no index streamed per slot), number of
 * patterns, whether int16 offsets fit, sliced-ELL slots */
int rsd_problem_layout(rsd_problem* pb, int64_t* out);
/* layout of the FP32 copy used by rsd_spmv_f32 and the mixed CG (no reference
 * counterpart; decided here on first call if no FP32 work came first):
 * out[4] = column mode (3: symmetric half storage, the upper value of each
 * symmetric pair kept once, below-diagonal entries read from the row above;
 * 4: the same in uniform stencil storage, one dense plane per non-negative
 * column offset, used when every row's offsets lie in one 27- (hex) or
 * 9-point (quad) stencil; else the FP64 layout's mode), number of patterns,
 * upper entries per row (half storage) or planes (uniform stencil; 0
 * otherwise), FP32 slots */
int rsd_problem_layout_f32(rsd_problem* pb, int64_t* out);
/* Assembly scheme (no reference counterpart; both give bitwise the same
 * matrices and lifted vectors).  mode 1 (the default): row gather -- one
 * pass writes every element matrix to an element buffer (n_elem * npe^2
 * doubles, reserved on first use), a second pass has each row sum its
 * elements' rows in colour order and write its values once; the essential
 * lift likewise (the border elements' terms, then each affected row adds
 * them in colour order); used when the problem's gather tables exist
 * (elements around a node at most 16, rows at most 64 entries).  mode 0:
 * the element scatter, colour by colour (frees the element buffers).
 * *in_use (nullable) receives 1 when the row gather is in use. */
int rsd_problem_set_assembly(rsd_problem* pb, int32_t mode, int32_t* in_use);
/* Element arithmetic (no reference counterpart).  RSD_ARITH_REFERENCE (the
 * default): every element kernel follows the reference's chain of products
 * and sums, each rounded separately -- bitwise the CPU loops in the same
 * element order.  RSD_ARITH_CONTRACTED: the Backward Euler form pair
 * (isotropic diffusion plus scaled mass, state-free model) -- its matrix
 * with the mass vector, and its operand products (the matrix-free apply) --
 * forms per-point factors once and accumulates in fused multiply-adds, about
 * a quarter of the FP64 operations; entries then differ from the reference
 * order by a few ulps of the largest entry of their row.  RSD_ARITH_UNIFORM:
 * as contracted, and on a mesh whose elements are all translates of
 * element 0 (a block mesh; checked once, within 1e-10 of the element's
 * extent) with the row gather in use, the pair's element matrix is
 * kappa_e Kh + coef Mh with Kh and Mh formed once on element 0: the matrix
 * comes from the row pass alone (no element pass, no element buffer) and
 * the operand product from an 8 x 8 element product; values then differ
 * from the per-element reference by the rounding of each element's own
 * coordinates, about 1e-14 relative (measured within 3e-15); selecting it
 * reserves its buffers (2 n_elem npe doubles), so no later call allocates.
 * RSD_ARITH_FACTORED: as contracted, and on Q1 elements of any shape with
 * the row gather in use, the pair's operand product (the matrix-free apply,
 * FEMOperator.hpp:38-56) is formed without element matrices: per
 * quadrature point the geometry, the operand's value and gradient from the
 * element's multilinear coefficients (sum factorisation of the Q1 basis),
 * then the flux and mass terms tested in the same coefficients; element
 * vectors are summed by row in colour order.  Products differ from the
 * reference order by rounding, about 1e-15 of sum_j |A_ij x_j|; selecting
 * it reserves n_elem npe doubles.  Every other form and kernel keeps the
 * reference order. */
#define RSD_ARITH_REFERENCE 0
#define RSD_ARITH_CONTRACTED 1
#define RSD_ARITH_UNIFORM 2
#define RSD_ARITH_FACTORED 3
int rsd_problem_set_arithmetic(rsd_problem* pb, int32_t mode);
/* padding the uniform stencil kernel needs around its x vector (no reference
 * counterpart): rsd_spmv_f32_padded reads x[-lo .. n + hi - 1], which must
 * be finite (zeros); 0 and 0 when the layout is not the uniform stencil */
int rsd_problem_padding_f32(rsd_problem* pb, int64_t* lo, int64_t* hi);
/* Binary mesh files, PMSH v1 (src/io/MeshIO.cpp:31-140; magic 0x504D5348,
 * io/MeshIO.hpp:49-50).  Host-side, no device work.  rsd_pmsh_info:
 * out[12] = parametric dim, spatial dim, element family (0 quad, 2 hex),
 * basis type (0 Lagrange), n_nodes, n_elem, nodes per element, faces per
 * element, number of basis orders, orders[0..2].  rsd_pmsh_read fills host
 * xyz [n_nodes*sdim], ien [n_elem*npe], rng [n_elem*faces] (MeshIO::readBinary);
 * rsd_pmsh_write writes them (MeshIO::writeBinary, no extraction operators). */
int rsd_pmsh_info(const char* path, int64_t* out);
int rsd_pmsh_read(const char* path, double* xyz, int64_t* ien, int32_t* rng);
int rsd_pmsh_write(const char* path, int32_t pdim, int32_t sdim, int32_t family, int32_t basis_type,
                   int64_t n_nodes, int64_t n_elem, int32_t npe, int32_t faces, int32_t n_orders,
                   const int32_t* orders, const double* xyz, const int64_t* ien, const int32_t* rng);
/* a problem on the mesh of a PMSH file (Lagrange quadrilaterals in 2D,
 * hexahedra in 3D), as rsd_problem_create; HeatDispatcher.cpp:48 reads the
 * mesh file this way before building the problem */
int rsd_problem_create_pmsh(rsd_ctx* ctx, const char* path, int32_t nq, int32_t n_bc, const int32_t* bc_tags,
                            const char* const* bc_exprs, rsd_problem** out);
/* Nodal field files, PNDF v1 (io/fieldio/FieldIO.hpp:19-26,
 * FieldIO.tpp:126-231).  rsd_problem_nodal_field gives the host nodal field
 * [n_nodes] of the free device vector d_U, with the Dirichlet values at
 * `time` on constrained nodes (FieldIO::reconstructNodalField);
 * rsd_pndf_write writes it (FieldIO::writeBinary); rsd_pndf_read reads host
 * nodal [n_nodes] and the time, refusing a file whose mesh tag (FNV-1a over
 * xyz and ien, FieldIO::computeMeshTag) is not the problem's. */
int rsd_problem_nodal_field(rsd_ctx* ctx, rsd_problem* pb, double time, const double* d_U, double* nodal);
int rsd_pndf_write(rsd_ctx* ctx, rsd_problem* pb, double time, const double* d_U, const char* path);
int rsd_pndf_read(rsd_problem* pb, const char* path, double* nodal, double* time);
int rsd_problem_mesh_tag(rsd_problem* pb, uint64_t* out);
/* d_out [n_free] = f(t, x, y, z) at each free DOF's node: an initial
 * condition given by an expression (InitialConditionConfigParser.cpp with
 * read mode "expression"), same grammar as the boundary expressions */
int rsd_problem_eval_expr(rsd_ctx* ctx, rsd_problem* pb, const char* expr, double time, double* d_out);
/* Expressions on the host, in the grammar the device evaluates (the
 * Dirichlet values, loads and k(T) above): the host side of a function that
 * also runs on the device, e.g. a model evaluated by the reference's CPU
 * assembly.  Replaces the exprtk-backed ScalarExpression /
 * TemperatureExpression call (src/utils/expression/ScalarExpression.cpp:3-30,
 * equation/heateq/evaluator/TemperatureExpression.hpp:38-51); variables t x y
 * z T.  rsd_expr_uses tells whether the expression reads t and T. */
int rsd_expr_compile(const char* text, rsd_expr** out);
int rsd_expr_eval(const rsd_expr* e, double t, double x, double y, double z, double T, double* out);
int rsd_expr_uses(const rsd_expr* e, int32_t* uses_t, int32_t* uses_T);
int rsd_expr_destroy(rsd_expr* e);
/* Dirichlet values given by the caller instead of the problem's expressions
 * (gatherElementVector's Full / Constrained modes evaluate the registered
 * boundary functions at the constrained nodes, ElementMap.tpp:39-72): host
 * values [n_nodes - n_free] at the constrained nodes in ascending node
 * order, valid for `time` only -- an assembly at another time fails with
 * RSD_ERR_ARG until values for it are given.  values == NULL returns the
 * problem to its expressions. */
int rsd_problem_set_dirichlet_values(rsd_ctx* ctx, rsd_problem* pb, double time, const double* values);
/* topo->alg map (-1 constrained), alg->topo, constraint tag per node */
int rsd_problem_get_dofs(rsd_problem* pb, int32_t* topo_to_alg, int32_t* alg_to_topo, int32_t* ctag);
/* CSR pattern as the reference's size_t arrays, widened to int64 */
int rsd_problem_get_pattern(rsd_problem* pb, int64_t* row_ptr, int64_t* col_idx);
/* element -> nnz map in CSR positions (-1 constrained), [ne*npe*npe] */
int rsd_problem_get_elem_map(rsd_problem* pb, int64_t* map);
/* element colour and colour-major assembly order */
int rsd_problem_get_colours(rsd_problem* pb, int32_t* colour, int64_t* order);

/* conductivity K = kappa_e * k(T) * shape (ConductivityModel.hpp:60-96; the
 * per-element factor kappa_e has no reference counterpart).  k_expr / dk_expr
 * nullable (constant k_scalar, k' = 0); tensor nullable (isotropic), else
 * dim*dim row-major; kappa nullable, else host [n_elem]. */
int rsd_problem_set_conductivity(rsd_problem* pb, double k_scalar, const char* k_expr,
                                 const char* dk_expr, const double* tensor, const double* kappa);
/* DensityModel (DensityModel.hpp:13) and constant SpecificHeatModel */
int rsd_problem_set_mass(rsd_problem* pb, double rho, double cp);
/* per-element kappa from device memory [n_elem] (adjoint sweeps) */
int rsd_problem_set_kappa_device(rsd_problem* pb, const double* d_kappa);

/* ---------------------------------------------------------------------- */
/* matrices on the problem's pattern: CSRMatrix<Real,Hip>                  */
/* (linalg/types/CSRMatrix.hpp:14-88).  Values live in HBM in a sliced     */
/* ELL layout (64-row slices, one wavefront each) for coalesced SpMV.      */
/* ---------------------------------------------------------------------- */
int rsd_matrix_create(rsd_problem* pb, rsd_matrix** out);
int rsd_matrix_destroy(rsd_matrix* m);
int rsd_matrix_zero(rsd_ctx* ctx, rsd_matrix* m);                        /* CSRMatrix::zero */
/* values in CSR order, host [nnz] */
int rsd_matrix_get_values(rsd_ctx* ctx, rsd_matrix* m, double* csr_values);
int rsd_matrix_set_values(rsd_ctx* ctx, rsd_matrix* m, const double* csr_values);
/* the values in CSR order into device memory [nnz] (a CSRMatrix<Real,Hip>'s
 * data(), CSRMatrix.hpp:76-78); rsd_matrix_load_csr below is the reverse */
int rsd_matrix_store_csr(rsd_ctx* ctx, rsd_matrix* m, double* d_csr_values);

/* y = A x: operations::matvec (linalg/operations/backend/cpu/MatrixOps.tpp:5-23),
 * row sums in the reference's left-to-right order (bitwise equal) */
int rsd_spmv_f64(rsd_ctx* ctx, rsd_matrix* A, const double* x, double* y);
/* The FP64 stencil copy (no reference counterpart; the product is
 * operations::matvec, MatrixOps.tpp:5-23, bitwise).  When every row's column
 * offsets lie in one 27-point (hex Q1) or 9-point (quad Q1) stencil of runs
 * of three consecutive offsets (block meshes), a second copy of the values
 * holds all stencil positions of each 64-row slice, zeros where a row lacks
 * one, in 16-byte pairs; its product loads one x entry per run and lane and
 * passes the neighbours between lanes.  rsd_cg_solve_f64, rsd_be_run and the
 * other FP64 CG solves refresh and use it by themselves.
 * rsd_problem_layout_stencil: out[2] = stencil points (0: no stencil copy),
 * slots of the copy.  rsd_matrix_to_stencil refreshes the copy from A's
 * values (RSD_ERR_ARG without a stencil); rsd_spmv_f64_stencil is y = A x
 * over it, bitwise rsd_spmv_f64's result. */
int rsd_problem_layout_stencil(rsd_problem* pb, int64_t* out);
int rsd_matrix_to_stencil(rsd_ctx* ctx, rsd_matrix* A);
int rsd_spmv_f64_stencil(rsd_ctx* ctx, rsd_matrix* A, const double* x, double* y);
/* The slice-union copy the GMRES products read (no reference counterpart;
 * the product is operations::matvec, MatrixOps.tpp:5-23, bitwise in FP64 and
 * the mixed solver's FP32 row sums in FP32).  When the rows share offset
 * patterns and every 64-row slice's union of its rows' column offsets has at
 * most 32 positions (Q1 and Q2 quads, Q1 hexes), a copy holds each slice's
 * union positions, zeros where a row lacks one, in 16-byte pairs, so the
 * rows of a slice share their offsets even where their own patterns differ.
 * rsd_gmres_solve_f64 and rsd_gmres_solve_mixed refresh and use it by
 * themselves.  out[2] = 1 when the problem has the layout (else 0), slots
 * of the copy. */
int rsd_problem_layout_union(rsd_problem* pb, int64_t* out);
/* Jacobi::update (linalg/solver/preconditioner/Jacobi.hpp:23-36) via
 * CSROperator::diagonal (CSROperator.hpp:40-45): 1/d, or 1 if |d| <= 1e-300 */
int rsd_jacobi_inv_diag(rsd_ctx* ctx, rsd_matrix* A, double* inv_diag);

/* ---------------------------------------------------------------------- */
/* solvers                                                                  */
/* ---------------------------------------------------------------------- */
/* SolverReport (linalg/solver/base/SolverReport.hpp:12-28) */
typedef struct rsd_report {
    int32_t converged;
    int64_t iterations;
    double initial_residual;
    double final_residual;
    double final_residual_rel;
} rsd_report;

/* tolerance type (cg/Config.hpp:12, gmres/Config.hpp:12 ToleranceType) */
#define RSD_TOL_RELATIVE 0 /* ||r_k|| / (||r_0|| + 1e-50) < tol */
#define RSD_TOL_ABSOLUTE 1 /* ||r_k|| < tol */

/* cg::Solver::solve (linalg/solver/iterative/cg/Solver.tpp:15-119): PCG with
 * the reference's stopping test (tol_type), warm start from x.  Fused
 * kernels, scalars kept on the device.  hist (host, nullable,
 * >= max_iters entries) receives ||r_k||. */
int rsd_cg_solve_f64(rsd_ctx* ctx, rsd_matrix* A, const double* b, double* x, double tol, int32_t tol_type,
                     int64_t max_iters, int32_t precond, rsd_report* rep, double* hist);

/* gmres::Solver::solve (linalg/solver/iterative/gmres/Solver.tpp:5-170):
 * right-preconditioned restarted GMRES(krylov) with modified Gram-Schmidt,
 * Givens rotations, breakdown at 1e-14 and the true residual at every
 * restart; rep->iterations counts Arnoldi steps. */
int rsd_gmres_solve_f64(rsd_ctx* ctx, rsd_matrix* A, const double* b, double* x, double tol, int32_t tol_type,
                        int64_t max_iters, int32_t krylov, int32_t precond, rsd_report* rep);
/* Mixed-precision GMRES (no reference counterpart; the reference is FP64
 * only): FP64 iterative refinement -- r = b - A x with the FP64 values,
 * stopping on ||r|| / (||r_0|| + 1e-50) < tol, x += ||r|| d -- around an
 * inner right-preconditioned GMRES(krylov) on A d = r / ||r|| with a full
 * FP32 copy of A's values (any matrix, symmetric or not), FP32 basis and
 * products, Gram-Schmidt and norm dots summed in FP64, Hessenberg and Givens
 * in FP64.  The inner relative tolerance is max(inner_floor, 0.5 tol
 * ||r_0|| / ||r||), capped at 0.5 (inner_floor <= 0: 1e-4).
 * rep->iterations counts inner Arnoldi steps; outer_cycles (nullable) the
 * refinements. */
int rsd_gmres_solve_mixed(rsd_ctx* ctx, rsd_matrix* A, const double* b, double* x, double tol, int64_t max_iters,
                          int32_t krylov, int32_t precond, double inner_floor, rsd_report* rep,
                          int64_t* outer_cycles);

/* ---------------------------------------------------------------------- */
/* raw device CSR in the reference's layout (linalg/types/CSRMatrix.hpp:14-88:
 * size_t row pointer [n+1], size_t columns [nnz], values [nnz]), all
 * pointers device memory.  Row sums run left to right in CSR order.       */
/* ---------------------------------------------------------------------- */
/* operations::matvec (linalg/operations/backend/cpu/MatrixOps.tpp:5-23) */
int rsd_csr_spmv_f64(rsd_ctx* ctx, int64_t n_rows, const uint64_t* row_ptr, const uint64_t* col_idx,
                     const double* val, const double* x, double* y);
/* operations::lump (MatrixOps.tpp:26-44): d_i = sum of row i */
int rsd_csr_lump_f64(rsd_ctx* ctx, int64_t n_rows, const uint64_t* row_ptr, const double* val, double* d);
/* CSROperator::diagonal (linalg/operator/CSROperator.hpp:40-45) with
 * CSRMatrix::getDataIndex (CSRMatrix.hpp:49-73): a row without a diagonal
 * entry fails with RSD_ERR_ARG, as the reference throws std::out_of_range */
int rsd_csr_diag_f64(rsd_ctx* ctx, int64_t n_rows, const uint64_t* row_ptr, const uint64_t* col_idx,
                     const double* val, double* d);
/* Jacobi::update (linalg/solver/preconditioner/Jacobi.hpp:33-36):
 * inv_i = 1/d_i where |d_i| > 1e-300, else 1; d == inv is allowed */
int rsd_inv_diag_f64(rsd_ctx* ctx, int64_t n, const double* d, double* inv);
/* a solver matrix (sliced ELL) on the pattern of a square device CSR with
 * sorted unique columns, for the fused solvers above; values are loaded in
 * CSR order with rsd_matrix_load_csr.  Replaces the CSRMatrix the
 * reference's cg / gmres Solver templates read through CSROperator. */
int rsd_matrix_from_csr(rsd_ctx* ctx, int64_t n_rows, const uint64_t* row_ptr, const uint64_t* col_idx,
                        rsd_matrix** out);
int rsd_matrix_load_csr(rsd_ctx* ctx, rsd_matrix* A, const double* val);
/* (rsd_matrix_load_csr also loads a problem's matrix, rsd_matrix_create,
 * from device values in the problem's CSR order) */

/* Mixed precision (no reference counterpart; the reference is FP64 only).
 * rsd_matrix_to_f32 refreshes the FP32 copy of A's values; rsd_spmv_f32 is
 * y = A x with FP32 values and FP32 accumulation.  When the pattern is
 * structurally symmetric and its rows share offset patterns (block meshes),
 * the copy keeps each symmetric pair once, the upper value, and the product
 * uses it for both entries: the symmetric operator the mixed CG needs (an
 * assembled heat matrix is symmetric to FP64 rounding; the conversion flags a
 * pair that differs by more than 1e-10 of its row's largest entry, and
 * rsd_cg_solve_mixed then fails with RSD_ERR_ARG).  rsd_problem_layout_f32
 * tells which layout is in use.  rsd_cg_solve_mixed: iterative refinement
 * with an FP64 outer residual r = b - A x and an FP32 inner PCG whose dot
 * products accumulate in FP64; stops on the FP64 relative residual
 * ||r|| / (||r_0|| + 1e-50) < tol.  inner_floor bounds the inner relative
 * tolerance from below (<= 0: 1e-4).  rep->iterations counts inner
 * iterations (SpMVs). */
int rsd_matrix_to_f32(rsd_ctx* ctx, rsd_matrix* A);
int rsd_spmv_f32(rsd_ctx* ctx, rsd_matrix* A, const float* x, float* y);
/* rsd_spmv_f32 by the mixed CG's own kernel: with the uniform stencil
 * storage every row reads the same 27 (9) entries, x three at a time, with no
 * per-row pattern, reading x at rsd_problem_padding_f32's padding (which the
 * caller provides, zeroed); bitwise rsd_spmv_f32's result.  Other layouts:
 * rsd_spmv_f32. */
int rsd_spmv_f32_padded(rsd_ctx* ctx, rsd_matrix* A, const float* x, float* y);
int rsd_cg_solve_mixed(rsd_ctx* ctx, rsd_matrix* A, const double* b, double* x, double tol, int64_t max_iters,
                       int32_t precond, double inner_floor, rsd_report* rep, int64_t* outer_cycles);

/* ---------------------------------------------------------------------- */
/* assembly (fem/assembly/backend/cpu/Assembler.tpp:86-255,                */
/* fem/boundary/backend/cpu/BoundaryApplicator.tpp:7-64).  Element kernels  */
/* run colour by colour: elements of one colour share no node, so the       */
/* scatter through the element->nnz map needs no atomics and the result is  */
/* deterministic.  U / Udot / operand are free-DOF device vectors.         */
/* ---------------------------------------------------------------------- */
int rsd_assemble_matrix(rsd_ctx* ctx, rsd_problem* pb, double time, int32_t n_forms,
                        const int32_t* kinds, const double* coefs, const double* U,
                        const double* Udot, rsd_matrix* K);
int rsd_assemble_vector(rsd_ctx* ctx, rsd_problem* pb, double time, int32_t n_forms,
                        const int32_t* kinds, const double* coefs, const double* U,
                        const double* Udot, const double* operand, int32_t selector, double* F);
/* SourceForm (SourceForm.hpp:22-32) with expression f(t,x,y,z) (nullable = 0)
 * or NodalSourceForm (NodalSourceForm.hpp:62-86) with device nodal [n_nodes] */
int rsd_assemble_load(rsd_ctx* ctx, rsd_problem* pb, double time, const char* src_expr,
                      const double* nodal, double* F);
/* essential lift F -= A_fc g (BoundaryApplicator.tpp:7-64) */
int rsd_apply_essential(rsd_ctx* ctx, rsd_problem* pb, double time, int32_t n_forms,
                        const int32_t* kinds, const double* coefs, double* F);
/* Natural (flux) boundary conditions: BoundaryApplicator::applyNaturalBCs
 * (fem/boundary/backend/cpu/BoundaryApplicator.tpp:66-140) with
 * FluxBoundaryForm (equation/heateq/form/FluxBoundaryForm.hpp:11-38) or
 * NodalFluxForm (NodalFluxForm.hpp:12-55).  F += for every element
 * (ascending) and face (ascending) whose tag has entries, the face vector
 * Fe[a] = sum over the face's Gauss points (nq per direction), the tag's
 * entries in order and sD of (q_sD N_a) n_sD w, n the unnormalised cofactor
 * normal; added at the face's free nodes in that order, bitwise the serial
 * loop.  Entry k (at most 8) applies to faces tagged entry_tags[k]; its flux
 * vector q is the nodal field d_nodal_flux[k] (device [n_nodes * dim],
 * interpolated with the face shape values) when d_nodal_flux and that
 * pointer are non-null, else the dim expressions flux_exprs[k * dim + sD]
 * of (t, x, y, z). */
int rsd_apply_natural(rsd_ctx* ctx, rsd_problem* pb, double time, int32_t n_entries, const int32_t* entry_tags,
                      const char* const* flux_exprs, const double* const* d_nodal_flux, double* F);
/* Assembler::assembleDiagonal (Assembler.tpp:142-197): diag [n_free] zeroed,
 * then each element's Ke[a][a] added in colour order; bitwise the diagonal
 * of rsd_assemble_matrix with the same forms.  FEMOperator::diagonal. */
int rsd_assemble_diagonal(rsd_ctx* ctx, rsd_problem* pb, double time, int32_t n_forms,
                          const int32_t* kinds, const double* coefs, const double* U, double* diag);

/* Matrix-free operator (linalg/operator/FEMOperator.hpp:27-71):
 * FEMOperator::apply is rsd_assemble_vector with selector RSD_SEL_OPERAND
 * (y = forms applied to the free vector x, element by element).
 * rsd_cg_solve_fem is cg::Solver::solve (cg/Solver.tpp:15-119) over that
 * operator, Jacobi from rsd_assemble_diagonal: no matrix is formed.  U / Udot
 * are the state the forms see (nullable); arguments otherwise as
 * rsd_cg_solve_f64. */
int rsd_cg_solve_fem(rsd_ctx* ctx, rsd_problem* pb, double time, int32_t n_forms,
                     const int32_t* kinds, const double* coefs, const double* U, const double* Udot,
                     const double* b, double* x, double tol, int32_t tol_type, int64_t max_iters,
                     int32_t precond, rsd_report* rep, double* hist);

/* Weak boundary reaction per tag: QuantityEvaluator<CPU>::
 * evaluateBoundaryReaction (fem/quantity/backend/cpu/QuantityEvaluator.tpp:
 * 95-199) with the heat problem's pairs (HeatProblem.tpp:385-410): source
 * src_expr (nullable) on the value, -1 x stiffness on the value, -1 x mass on
 * d_Udot (nullable: steady).  d_U / d_Udot are device free vectors [n_free];
 * Dirichlet nodes take their values at `time`.  result[t] (host) receives
 * the sum of the element residual over the nodes of the faces tagged
 * tags[t], measure[t] (host) the faces' measure, both bitwise the reference's
 * serial sums (BoundaryQuantityRegistry::result / measureSum,
 * BoundaryQuantityRegistry.hpp:17-50).  The average a monitor reports is
 * result / measure.  The plan for a tag list is built on first use and
 * cached on the problem. */
int rsd_boundary_reaction(rsd_ctx* ctx, rsd_problem* pb, double time, const char* src_expr,
                          const double* d_U, const double* d_Udot, int32_t n_tags,
                          const int32_t* tags, double* result, double* measure);
/* Strong (pointwise) boundary heat flux per tag: QuantityEvaluator<CPU>::
 * evaluateBoundaryRegistry (fem/quantity/backend/cpu/QuantityEvaluator.tpp:
 * 14-93) with HeatFluxIntegrand (equation/heateq/quantity/
 * HeatFluxIntegrand.hpp:19-38), Integral reduction: result[t] (host) = the
 * sum over the points of the faces tagged tags[t] of -(K gradT).n w (n the
 * unnormalised cofactor normal, T the full gather of d_U at `time`, K =
 * k(T) x shape), measure[t] (host) the faces' measure, both bitwise the
 * reference's serial sums.  The Average reduction (QuantityForms::finalize)
 * is result / measure. */
int rsd_boundary_flux(rsd_ctx* ctx, rsd_problem* pb, double time, const double* d_U, int32_t n_tags,
                      const int32_t* tags, double* result, double* measure);

/* ---------------------------------------------------------------------- */
/* drivers                                                                  */
/* ---------------------------------------------------------------------- */
/* Linear Backward Euler on the problem: BackwardEulerRunner::step
 * (solver/timestepper/BackwardEuler.hpp:32-72) over BackwardEulerStage
 * (solver/stage/BackwardEulerStage.hpp:71-163) and Transient::solve
 * (solver/driver/Transient.hpp:21-35), CSR operator, CG + preconditioner.
 * d_U holds the initial free vector and receives the final one.  src_expr
 * nullable.  iters (host, nullable, >= max_steps) receives CG iterations per
 * step.  use_graph != 0 replays each step's work from hipGraphs.  precision 0:
 * FP64 PCG (bitwise reproducible); 1: mixed precision (rsd_cg_solve_mixed,
 * iters counts inner iterations).  Returns the number of steps in *steps. */
#define RSD_PREC_FP64 0
#define RSD_PREC_MIXED 1
int rsd_be_run(rsd_ctx* ctx, rsd_problem* pb, double t0, double tf, double dt,
               const char* src_expr, double cg_tol, int64_t cg_max_iters, int32_t precond,
               int32_t precision, int32_t use_graph, double* d_U, int64_t max_steps, int64_t* steps,
               int64_t* iters, double* timing_ms);
/* timing_ms (host, nullable, [2]): GPU time of the assembly phases and of
 * the solve phases summed over the steps, from HIP events on the stream. */

/* Steady nonlinear solve: NewtonRunner::solve (solver/nonlinear/Newton.hpp:
 * 32-83) over SteadyStage in nonlinear mode (solver/stage/SteadyStage.hpp:
 * 60-116): per iteration assemble the load, R = F - K(U) U with Dirichlet
 * values gathered in, ||R||, stop if ||R|| < abs_tol or ||R||/||R_0|| <
 * rel_tol, else reassemble J = K + K_T in place, dU = 0, solve J dU = R
 * (lin_solver 0 CG, 1 GMRES(krylov), 2 rsd_gmres_solve_mixed's
 * mixed-precision GMRES(krylov), 3 rsd_cg_solve_mixed's mixed-precision CG
 * on a full FP32 copy of J, which need not be symmetric), U += dU.  newton_iters receives the
 * iteration at which it converged (-1 not converged); res_hist (host,
 * nullable, >= max_newton) receives ||R||, lin_iters (host, nullable) the
 * inner iterations per Newton step. */
int rsd_newton_run(rsd_ctx* ctx, rsd_problem* pb, const char* src_expr, double abs_tol, double rel_tol,
                   int64_t max_newton, int32_t lin_solver, double lin_tol, int64_t lin_max_iters,
                   int32_t krylov, int32_t precond, double* d_U, int64_t* newton_iters, double* res_hist,
                   int64_t* lin_iters);

/* Adjoint gradient (no reference counterpart).  Loss 1/2 sum_k (U_N[obs_k]
 * - t_obs_k)^2 after n_steps linear Backward Euler steps of fixed dt from
 * d_U0 (free vector) with per-element conductivity d_kappa = exp(theta)
 * (device [n_elem]); forward and adjoint solves are Jacobi-PCG to cg_tol.
 * Every U_k is checkpointed in HBM.  want_grad != 0 writes dL/dtheta to
 * d_grad (device [n_elem]); d_UN (device, nullable) receives U_N.  obs and
 * t_obs are host arrays of free-DOF indices and values. */
int rsd_adjoint_grad_logk(rsd_ctx* ctx, rsd_problem* pb, const double* d_kappa, double t0, double dt,
                          int64_t n_steps, double cg_tol, int64_t cg_max_iters, const double* d_U0,
                          int64_t n_obs, const int64_t* obs, const double* t_obs, int32_t want_grad,
                          double* loss, double* d_grad, double* d_UN);

/* Batched PCG (no reference counterpart): many independent systems solved
 * together.  The batch holds the matrices (not copies; they must outlive
 * it); vectors are concatenated per problem at the offsets returned by
 * rsd_batch_layout (each segment padded to a multiple of 256 rows; padding
 * entries are ignored).  Each problem follows cg::Solver::solve
 * (cg/Solver.tpp:15-119) and its result is bitwise identical to
 * rsd_cg_solve_f64 on that problem alone.  When every problem has an FP64
 * stencil copy of the same size (rsd_problem_layout_stencil), the products
 * read those copies, which each call refreshes from the values.  reps: host
 * [n_problems]. */
int rsd_batch_create(rsd_ctx* ctx, int32_t n_problems, rsd_matrix* const* mats, rsd_batch** out);
int rsd_batch_destroy(rsd_batch* b);
int rsd_batch_layout(rsd_batch* b, int64_t* total_rows, int64_t* offsets);
int rsd_cg_solve_batched(rsd_ctx* ctx, rsd_batch* b, const double* rhs, double* x, double tol, int64_t max_iters,
                         int32_t precond, int32_t use_graph, rsd_report* reps);
/* Batched mixed precision (no reference counterpart): every problem of the
 * batch runs rsd_cg_solve_mixed -- FP32 inner CG on the FP32 copy (refreshed
 * from each matrix's values by every call), FP64 outer residual -- in one
 * set of launches; each problem's result, report and outer cycle count
 * (outer_cycles: host [n_problems], nullable) are bitwise those of
 * rsd_cg_solve_mixed on that problem alone.  Either every problem's FP32
 * copy has the symmetric half storage (rsd_problem_layout_f32), or none has
 * (RSD_ERR_ARG otherwise); when every problem is on a uniform stencil, the
 * products are the uniform-stencil kernel of the standalone solve.
 * RSD_ERR_ARG names a problem whose matrix fails the FP32 copy's symmetry
 * check.  The batch's schedule
 * applies (chunk: inner iterations per graph launch). */
int rsd_cg_solve_batched_mixed(rsd_ctx* ctx, rsd_batch* b, const double* rhs, double* x, double tol,
                               int64_t max_iters, int32_t precond, double inner_floor, rsd_report* reps,
                               int64_t* outer_cycles);
/* The batch's schedule (no reference counterpart): sched[4] = iterations per
 * chunk launch (1..256), row-product load batch (4 or 9), block order (0
 * problem by problem, 1 interleaved across problems), XCD block remap (0/1).
 * No schedule changes an operation or its order: results are bitwise the
 * same under every one.  Default 8, 9, 0, 1.  rsd_batch_autotune times the
 * 32 combinations of {4, 8, 16, 32} x {4, 9} x {0, 1} x {0, 1} on the batch
 * itself (about `iters` iterations each from b = 1, x = 0, in its own
 * scratch vectors, with the chunk launch, state read and synchronisation of
 * a solve), keeps the fastest and returns it in sched (nullable) with its
 * time per iteration in us_per_iter (nullable).  cache_path (nullable): a
 * JSON object of schedules keyed by device architecture, problem count,
 * rows, nonzeros, column mode and preconditioner; a hit is used without
 * timing (*from_cache = 1), a miss is timed and written back. */
int rsd_batch_set_schedule(rsd_batch* b, const int32_t* sched);
int rsd_batch_get_schedule(rsd_batch* b, int32_t* sched);
int rsd_batch_autotune(rsd_ctx* ctx, rsd_batch* b, int32_t precond, const char* cache_path, int64_t iters,
                       int32_t* sched, double* us_per_iter, int32_t* from_cache);

/* ---------------------------------------------------------------------- */
/* seeded synthetic inputs (no reference counterpart)                       */
/* ---------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon et al., SC'11), counter (i, stream), key seed;
 * normal: Box-Muller on two 53-bit uniforms, uniform: the first of them.
 * Host output [n]. */
int rsd_philox_normal(uint64_t seed, uint32_t stream, int64_t n, double* out);
int rsd_philox_uniform(uint64_t seed, uint32_t stream, int64_t n, double* out);

#ifdef __cplusplus
}
#endif
#endif
