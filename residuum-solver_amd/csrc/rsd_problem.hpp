// rsd_problem.hpp -- the device-resident discretised heat problem and the
// sliced-ELL matrix layout shared by assembly, SpMV and the solvers.
#pragma once

#include "rsd_internal.hpp"
#include "rsd_rowdot.hpp"

#include <unordered_set>

namespace rsd {

// Material model passed by value to the element kernels.
struct Model {
    double k_scalar = 1.0;
    double tensor[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    int32_t aniso = 0;
    int32_t has_k = 0, has_dk = 0;
    double rho = 1.0, cp = 1.0;
    Expr k, dk;
};

// Sliced ELL ("SELL-64") layout of the problem's pattern.  Row r lives in
// slice s = r / 64, lane l = r % 64; its j-th entry (CSR order) sits at slot
// slice_ptr[s] + 64*j + l.  One wavefront walks one slice with perfectly
// coalesced value / column loads, and for stencil-like FEM rows the gathered
// x entries of neighbouring lanes are neighbours too.
struct Sell {
    int64_t n_rows = 0, n_slices = 0, slots = 0;
    std::vector<int64_t> h_slice_ptr;
    DBuf<int64_t> slice_ptr;   // [n_slices + 1]
    DBuf<int32_t> row_ptr;     // CSR row pointer (int32), gives row lengths
    DBuf<int32_t> col;         // [slots] absolute column (padding: own row)
    DBuf<int16_t> off16;       // [slots] column - row, when every offset fits
    DBuf<int32_t> diag;        // [n_rows] slot of the diagonal entry, -1 if absent
    bool has_off16 = false;
    // shared column-offset patterns (COL_PAT): pattern id per row and the
    // offset lists, when at most 255 distinct lists cover every row
    DBuf<uint8_t> pat_id;      // [n_rows]
    DBuf<int32_t> pat_off;     // [n_pat * pat_stride] column - row, CSR order
    DBuf<int32_t> pat_len;     // [n_pat]
    int32_t n_pat = 0, pat_stride = 0;
    bool has_pat = false;
    // symmetric half storage of the FP32 copy (COL_HALF), built on first
    // use when the rows share offset patterns, the pattern is structurally
    // symmetric and at most 255 (offset, upper index) lists cover every row
    DBuf<uint8_t> half_id;     // [n_rows]
    DBuf<int32_t> half_off;    // [n_half * pat_stride] column - row, CSR order
    DBuf<int32_t> half_mir;    // [n_half * pat_stride] index m in the upper list of row min(r, c)
    DBuf<int32_t> half_val;    // [n_half * pat_stride] m * half_pad + min(c - r, 0)
    DBuf<int32_t> half_len;    // [n_half] entries
    DBuf<int32_t> half_low;    // [n_half] entries left of the diagonal
    DBuf<int32_t> half_dst;    // [n_half * pat_stride] value index - row of each upper entry (conversion)
    int32_t n_half = 0, half_width = 0;
    int64_t half_pad = 0;      // rows rounded up to 64: the stride between upper-value planes
    bool has_half = false, half_tried = false;
    // uniform stencil storage of the half values (COL_UNI): one dense plane
    // per non-negative offset of the union, planes uni_stride apart with row
    // r at uni_front + r, zeros where a row lacks the entry; the per-pattern
    // tables (half_val, half_dst) index the same storage
    bool has_uni = false;
    int32_t uni_nr = 0, uni_planes = 0;
    int32_t uni_roff[kUniMaxRuns] = {};
    int32_t uni_vidx[3 * kUniMaxRuns] = {};
    int64_t uni_stride = 0, uni_front = 0;
    int64_t uni_lo = 0, uni_hi = 0;  // padding an x vector needs before and after its n entries
    // stencil layout of the FP64 copy the CG reads (COL_ST): when the union
    // of every row's column offsets is st_np = 27 (hex Q1) or 9 (quad Q1)
    // offsets in runs of three consecutive ones, each 64-row slice holds all
    // st_np stencil positions of its rows, zeros where a row lacks one, in
    // 16-byte pairs; st_mask[r] has bit b set when row r holds position b
    bool has_st = false;
    int32_t st_np = 0;
    int32_t st_off[kStMaxPos] = {};  // column - row of each position, ascending
    DBuf<uint32_t> st_mask;          // [n_rows]
    int64_t st_slots() const { return n_slices * st_np * kSlice; }
    StCols st_cols() const {
        StCols c;
        c.mask = st_mask.p;
        c.np = st_np;
        for (int k = 0; k < st_np / 3; ++k) c.coff[k] = st_off[3 * k + 1];
        return c;
    }

    // slice-union layout of the copies the GMRES products read (COL_SU):
    // when every slice's union of column offsets has at most kSuMax
    // positions and the copy has at most 5/4 the sliced layout's slots
    bool has_su = false;
    int64_t su_slots = 0;
    DBuf<int64_t> su_ptr;   // [n_slices + 1]
    DBuf<int32_t> su_off;   // [n_slices * kSuMax]
    DBuf<int32_t> su_nu;    // [n_slices]
    DBuf<uint32_t> su_mask; // [n_rows]
    SuCols su_cols() const {
        SuCols c;
        c.ptr = su_ptr.p;
        c.off = su_off.p;
        c.nu = su_nu.p;
        c.mask = su_mask.p;
        c.n_slices = n_slices;
        return c;
    }

    Cols cols() const {
        Cols c;
        c.slice_ptr = slice_ptr.p;
        c.rp = row_ptr.p;
        c.col = col.p;
        c.off16 = off16.p;
        c.pid = pat_id.p;
        c.poff = pat_off.p;
        c.plen = pat_len.p;
        c.stride = pat_stride;
        return c;
    }
    // the cheapest column mode this layout supports
    int mode() const { return has_pat ? COL_PAT : has_off16 ? COL_OFF16 : COL_ABS; }
    // the FP32 copy: symmetric half storage when built, else the FP64 layout
    int mode32() const { return has_half ? COL_HALF : mode(); }
    // the mixed solver's own SpMV, whose x vector carries the padding
    int mode_padded() const { return has_uni ? COL_UNI : mode32(); }
    int64_t slots32() const {
        return has_uni ? uni_stride * uni_planes : has_half ? half_pad * half_width : slots;
    }
    Cols cols32() const {
        if (!has_half) return cols();
        Cols c = cols();
        c.pid = half_id.p;
        c.poff = half_off.p;
        c.plen = half_len.p;
        c.pval = half_val.p;
        c.stride = pat_stride;
        if (has_uni) {
            c.u_nr = uni_nr;
            for (int k = 0; k < kUniMaxRuns; ++k) c.u_roff[k] = uni_roff[k];
            for (int k = 0; k < 3 * kUniMaxRuns; ++k) c.u_vidx[k] = uni_vidx[k];
        }
        return c;
    }
};

// launch kern<MODE> for the layout's column mode
#define RSD_LAUNCH_COLS(mode, kern, grid, block, stream, ...)                                        \
    do {                                                                                            \
        switch (mode) {                                                                             \
        case ::rsd::COL_PAT: hipLaunchKernelGGL(kern<::rsd::COL_PAT>, grid, block, 0, stream, __VA_ARGS__); break;     \
        case ::rsd::COL_OFF16: hipLaunchKernelGGL(kern<::rsd::COL_OFF16>, grid, block, 0, stream, __VA_ARGS__); break; \
        default: hipLaunchKernelGGL(kern<::rsd::COL_ABS>, grid, block, 0, stream, __VA_ARGS__); break;                 \
        }                                                                                           \
    } while (0)

// the same for kernels on the FP32 copy, which may also use COL_HALF
#define RSD_LAUNCH_COLS32(mode, kern, grid, block, stream, ...)                                      \
    do {                                                                                            \
        if ((mode) == ::rsd::COL_UNI)                                                               \
            hipLaunchKernelGGL(kern<::rsd::COL_UNI>, grid, block, 0, stream, __VA_ARGS__);          \
        else if ((mode) == ::rsd::COL_HALF)                                                         \
            hipLaunchKernelGGL(kern<::rsd::COL_HALF>, grid, block, 0, stream, __VA_ARGS__);         \
        else                                                                                        \
            RSD_LAUNCH_COLS(mode, kern, grid, block, stream, __VA_ARGS__);                          \
    } while (0)

}  // namespace rsd

namespace rsd {
struct ReactionPlan;  // boundary-reaction plan, built on first use per tag list
struct NaturalPlan;   // natural boundary condition plan, built on first use per tag list
}

struct rsd_matrix;

struct rsd_problem {
    rsd_problem() = default;
    rsd_problem(const rsd_problem&) = delete;
    rsd_problem& operator=(const rsd_problem&) = delete;
    ~rsd_problem();
    rsd_ctx* ctx = nullptr;
    // the matrices made on this problem (rsd_matrix_create): destroying the
    // problem orphans them, and every entry point then refuses them
    std::unordered_set<rsd_matrix*> matrices;
    int32_t dim = 0, p = 1, nq = 2, npe = 0, nfaces = 0;
    int64_t n_nodes = 0, n_elem = 0, n_free = 0, nnz = 0;
    int32_t max_row = 0;
    // host copies (setup products, served by the getters)
    std::vector<double> h_xyz;
    std::vector<int32_t> h_ien, h_rng, h_t2a, h_a2t, h_ctag, h_colour;
    std::vector<int32_t> h_rp, h_ci;   // CSR pattern, int32
    std::vector<int32_t> h_order;      // colour-major element order
    std::vector<int64_t> colour_off;   // [n_colours + 1] into h_order
    std::vector<int64_t> bcolour_off;  // [n_colours + 1] into border
    // Dirichlet registry
    std::vector<int32_t> bc_tags;
    std::vector<rsd::Expr> bc_exprs;
    rsd::DBuf<int32_t> d_bc_tags;
    rsd::DBuf<rsd::Expr> d_bc_exprs;
    bool bc_time_dependent = false;
    double gval_time = 0.0;
    bool gval_valid = false;
    // Dirichlet values given by the caller for one time at a time
    // (rsd_problem_set_dirichlet_values) instead of the expressions
    bool gval_external = false;
    rsd::DBuf<int32_t> cnodes;    // [n_nodes - n_free] constrained nodes, ascending
    rsd::DBuf<double> cvals;      // their values, staged
    // device arrays
    rsd::DBuf<double> xyz;        // [n_nodes * dim]
    rsd::DBuf<int32_t> ien;       // [n_elem * npe]
    rsd::DBuf<int32_t> t2a;       // [n_nodes]
    rsd::DBuf<int32_t> ctag;      // [n_nodes]
    rsd::DBuf<int32_t> order;     // [n_elem] colour-major
    rsd::DBuf<int32_t> border;    // elements with a constrained node, colour-major
    rsd::DBuf<int32_t> map_sell;  // [n_elem * npe * npe] SELL slot or -1
    rsd::DBuf<int32_t> csr_slot;  // [nnz] SELL slot of each CSR entry (matrices made from a CSR)
    // row gather assembly: per free row (SELL order, 64-row slices) the row's
    // elements in colour order, at most one per colour; gat_m = 0 when the
    // tables are not built
    int32_t gat_m = 0;            // element slots per row
    bool gather_on = true;        // matrices assemble by row gather when the tables exist
    int32_t arith = 0;            // RSD_ARITH_REFERENCE, _CONTRACTED, _UNIFORM or _FACTORED (the BE form pair)
    int32_t uniform_geom = -1;    // every element a translate of element 0: -1 not yet checked, 0 no, 1 yes
    rsd::DBuf<double> uhat;       // [2 * npe * npe] Kh then Mh of element 0 (uniform geometry)
    rsd::DBuf<double> vbuf;       // [n_elem * npe] element vectors of the uniform pair, reserved on first use
    rsd::DBuf<int32_t> gat_el;    // [n_slices * gat_m * 64] e * npe + a, -1 past the row's elements
    rsd::DBuf<uint32_t> gat_k;    // [n_slices * gat_m * k4 * 64] row slot position of column b (byte b % 4 of
                                  // dword b / 4, k4 = ceil(npe / 4)); 0xff: constrained column
    rsd::DBuf<double> ebuf;       // [n_elem * npe * npe] element matrices, reserved on first use
    rsd::DBuf<double> fbuf;       // [n_elem * npe] element mass vectors, reserved on first use
    rsd::DBuf<double> ntab;       // [n_qp * (npe + 1)] shape values N_b and weight w of each Gauss point
    // lift by row gather: the free rows of the elements touching a
    // constrained node, each with its (border position k * npe + a) list in
    // border (colour-major) order
    int64_t lift_nrows = 0;
    rsd::DBuf<int32_t> lift_row;  // [lift_nrows]
    rsd::DBuf<int32_t> lift_off;  // [lift_nrows + 1]
    rsd::DBuf<int32_t> lift_idx;  // [lift_off[lift_nrows]]
    rsd::DBuf<double> lift_buf;   // [border elements * npe] element lift vectors
    rsd::DBuf<double> gval;       // [n_nodes] Dirichlet value (0 on free nodes)
    rsd::DBuf<double> kappa;      // [n_elem] or empty
    const double* kappa_ext = nullptr;  // device kappa supplied by the caller
    rsd::Sell sell;
    rsd::Model model;
    // caches for compiled source expressions
    std::string src_text;
    rsd::Expr src_expr;
    std::shared_ptr<rsd::ReactionPlan> reaction;
    std::shared_ptr<rsd::NaturalPlan> natural;

    const double* kappa_ptr() const { return kappa_ext ? kappa_ext : (kappa.n ? kappa.p : nullptr); }
    int n_colours() const { return (int)colour_off.size() - 1; }
};

struct rsd_matrix {
    rsd_problem* pb = nullptr;  // null once the problem is destroyed
    rsd_ctx* ctx = nullptr;     // the problem's context, for the matrix's own release
    std::unique_ptr<rsd_problem> own_pb;  // pattern-only problem of a matrix made from a CSR
    rsd::DBuf<double> val;    // [slots] SELL order
    rsd::DBuf<float> val32;   // [sell.slots32()] FP32 copy for the mixed-precision path
    rsd::DBuf<float> val32f;  // [sell.slots] full FP32 copy in the FP64 layout (mixed GMRES, any matrix)
    mutable rsd::DBuf<double> val_st;  // [sell.st_slots()] FP64 copy in the stencil layout, a cache the FP64 CG refreshes
    mutable rsd::DBuf<double> val_su;  // [sell.su_slots] FP64 slice-union copy, a cache the FP64 GMRES refreshes
    mutable rsd::DBuf<float> val_su32; // [sell.su_slots] FP32 slice-union copy, a cache the mixed GMRES refreshes
    rsd::DBuf<int32_t> asym;  // [1] set by the FP32 conversion when the half layout's symmetry check fails
};

namespace rsd {
// refresh the per-node Dirichlet values for `time` (no-op when cached)
void update_gval(rsd_problem* pb, double time);
// a problem matrix's values between the sliced layout and CSR order (device
// memory both): to_csr copies SELL -> CSR, else CSR -> SELL
void launch_sell_csr(const rsd_matrix* m, double* csr, bool to_csr);
// sliced ELL layout of pb->h_rp / h_ci, uploaded
void build_sell(rsd_problem* pb, const std::vector<int32_t>& rowlen);
// the symmetric half layout of the FP32 copy, tried once per problem (host
// work and uploads: call outside stream capture)
void build_half(rsd_problem* pb);
}
