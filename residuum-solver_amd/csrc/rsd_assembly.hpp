// rsd_assembly.hpp -- element-kernel arguments and host entry points.
#pragma once

#include "rsd_problem.hpp"

namespace rsd {

struct AsmArgs {
    const double* xyz;
    const int32_t* ien;
    const int32_t* t2a;
    const double* gval;
    const int32_t* order;
    const int32_t* map;
    const double* kappa;
    const double* U;
    const double* Udot;
    const double* operand;
    const double* nodal;
    double* val;
    double* F;
    double time;
    int32_t selector;
    int32_t nf;
    int32_t kinds[4];
    double coefs[4];
    int32_t need_state;
    const int32_t* stop;  // nullable: vector kernels return at once when *stop != 0
    const double* mvU;    // matrix kernels, nullable: also scatter (M_e Ue)[a] of mvU into mvF
    double* mvF;
    double* ebuf;         // matrix kernels writing element matrices (row gather assembly)
    double* fbuf;         // and, with mvU, the element mass vectors
    const double* ntab;   // Gauss point table: N_b (b < npe) and w per point, npe + 1 doubles a point
    const double* axx;    // row gather with the mass vector, nullable: also axo = A axx, each row's
    double* axo;          // product bitwise the FP64 row product (sequential, CSR order)
};

AsmArgs base_args(rsd_problem* pb);
void assemble_matrix(rsd_problem* pb, double time, int32_t nf, const int32_t* kinds, const double* coefs,
                     const double* U, const double* Udot, rsd_matrix* K);
void assemble_vector(rsd_problem* pb, double time, int32_t nf, const int32_t* kinds, const double* coefs,
                     const double* U, const double* Udot, const double* operand, int32_t selector, double* F,
                     const int32_t* stop = nullptr);
// Assembler::assembleDiagonal (Assembler.tpp:142-197): diag zeroed, then
// each element's Ke[a][a] added in colour order
void assemble_diagonal(rsd_problem* pb, double time, int32_t nf, const int32_t* kinds, const double* coefs,
                       const double* U, double* diag);
// assemble_matrix plus, from the same element pass, the unit mass form's
// vector on Um (assemble_vector with one MASS form and RSD_SEL_VALUE) into
// Fm, bitwise both separate assemblies.  The Dirichlet values of Um's
// gather are those at time_m; the matrix's state (when its forms need one)
// uses those at `time`, so the two times must give the same values.
void assemble_matrix_mass(rsd_problem* pb, double time, int32_t nf, const int32_t* kinds, const double* coefs,
                          const double* U, const double* Udot, rsd_matrix* K, const double* Um, double time_m,
                          double* Fm, const double* axx = nullptr, double* axo = nullptr);
// whether assemble_matrix_mass can also form axo = K axx (the row gather pass)
bool gather_forms_product(const rsd_problem* pb);
// whether assemble_matrix_mass applies: one set of Dirichlet values serves
// both gathers
bool matrix_mass_fusable(const rsd_problem* pb, int32_t nf, const int32_t* kinds);
// reserves the element buffers of the row gather assembly (call before a
// stream capture that assembles matrices); with_mass: also the element mass
// vectors of assemble_matrix_mass
void gather_reserve(rsd_problem* pb, bool with_mass);
// what assembling these forms (with the mass vector) will allocate, reserved
// outside any stream capture
void assembly_reserve(rsd_problem* pb, int32_t nf, const int32_t* kinds, bool with_mass);
// the buffers of the problem's element arithmetic (uniform geometry:
// tables and element vectors; factored: element vectors), when the mesh and
// tables allow that arithmetic
void arith_reserve(rsd_problem* pb);
// the Gauss point table (shape values and weights), computed on the device
// by the element kernels' own functions; part of problem set-up
void build_shape_table(rsd_problem* pb);
void assemble_load(rsd_problem* pb, double time, const Expr* src, const double* nodal, double* F);
void apply_essential(rsd_problem* pb, double time, int32_t nf, const int32_t* kinds, const double* coefs, double* F);
const Expr* cached_source(rsd_problem* pb, const char* text);
void adjoint_grad_step(rsd_problem* pb, double time, const double* U, const double* lam, const double* kappa,
                       double* grad);

}  // namespace rsd
