"""GPU parity of the core path: setup products, SpMV, reductions, assembly,
CG and the Backward Euler loop, checked against the CPU oracle.

The oracle runs in GPU order (colour-major element loop, the kernels'
block-tree reductions), so everything here must agree bit for bit; the
reference-order oracle (itself bitwise equal to the reference's own code,
see test_oracle.py) bounds the rounding difference.
"""
import os
import sys

import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu

CASES = [
    # dim, n, p, nq, tags, exprs, ext
    (2, (4, 4), (1, 1), 2, [0, 1, 3], ["x"] * 3, (0.0, 4.0, 0.0, 4.0)),
    (2, (7, 5), (2, 2), 3, [0, 1, 2, 3], ["(1-x)*(1-y)"] * 4, None),
    (3, (5, 4, 3), (1, 1, 1), 2, [0, 1], ["1", "0"], None),
    (3, (6, 6, 6), (1, 1, 1), 2, [4], ["1"], None),
]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def build(ctx, case, kappa_seed=None):
    import residuum_hip as rh
    dim, n, p, nq, tags, exprs, ext = case
    say("case", case)
    pb = rh.Problem(ctx, dim, n, p, nq, tags, exprs, ext=ext)
    oc = O.Case(dim, n, p, nq, tags, exprs, ext=ext, order=O.ORDER_GPU)
    if kappa_seed is not None:
        kap = np.exp(0.5 * O.normal_field(kappa_seed, 0, oc.n_elem))
        oc.kappa = kap
        oc.set_order(O.ORDER_GPU)
        pb.set_conductivity(1.0, kappa=kap)
    return pb, oc


def same(a, b):
    return a.shape == b.shape and np.array_equal(a, b)


@pytest.mark.parametrize("case", CASES)
def test_setup_products_bitexact(ctx, case):
    pb, oc = build(ctx, case)
    xyz, ien, rng = pb.mesh()
    assert same(xyz, oc.xyz) and same(ien, oc.ien) and same(rng, oc.rng)
    t2a, a2t, ct = pb.dofs()
    assert pb.n_free == oc.n_free
    assert same(t2a, oc.t2a) and same(a2t, oc.a2t) and same(ct, oc.ctag)
    rp, ci = pb.pattern()
    assert same(rp, oc.rp) and same(ci, oc.ci)
    assert same(pb.elem_map(), oc.map)
    col, order = pb.colours()
    assert same(col, oc.colour) and same(order, oc.elem_order)


@pytest.mark.parametrize("case", CASES)
def test_spmv_bitexact(ctx, case):
    pb, oc = build(ctx, case)
    rng = np.random.default_rng(1)
    val = rng.standard_normal(oc.nnz)
    x = rng.standard_normal(oc.n_free)
    K = pb.matrix()
    K.set_values(val)
    assert same(K.values(), val)
    dx, dy = ctx.from_numpy(x), pb.vector()
    K.spmv(dx, dy)
    assert same(dy.numpy(), oc.spmv(val, x))


STENCIL_CASES = CASES + [
    # one element across in y and z: every free row sits on the boundary
    (3, (9, 1, 1), (1, 1, 1), 2, [0], ["1"], None),
    # 70 x 3 quads, all sides Dirichlet: rows whose stencils reach past 64 lanes
    (2, (70, 3), (1, 1), 2, [0, 1, 2, 3], ["0"] * 4, None),
    # 40 x 40 x 3 hexes, x faces Dirichlet: slices mixing line ends, a partial last slice
    (3, (40, 40, 3), (1, 1, 1), 2, [0, 1], ["1", "0"], None),
]


@pytest.mark.parametrize("case", STENCIL_CASES)
def test_spmv_stencil_bitexact(ctx, case):
    """The stencil copy's product (one x load per run and lane, neighbours by
    lane shifts) equals the sliced-ELL product bit for bit, also where x holds
    infinities and NaNs: positions a row lacks never reach its sum.  (A NaN
    result's sign is not fixed by IEEE 754 and follows the operand order the
    compiler picks for each add, so NaNs are compared by position.)"""
    import residuum_hip as rh
    pb, oc = build(ctx, case)
    npts, slots = pb.layout_stencil()
    K = pb.matrix()
    rng = np.random.default_rng(7)
    K.set_values(rng.standard_normal(oc.nnz))
    # the copy exists when the union of the rows' column offsets (oracle
    # pattern) is 9 or 27 offsets in runs of three consecutive ones
    rows = np.repeat(np.arange(oc.n_free), np.diff(oc.rp))
    uni = np.unique(oc.ci - rows)
    want = len(uni) in (9, 27) and all(np.array_equal(uni[k:k + 3], uni[k] + np.arange(3))
                                       for k in range(0, len(uni), 3))
    if not want:
        assert (npts, slots) == (0, 0)
        with pytest.raises(rh.RsdError):
            K.to_stencil()
        return
    assert npts == len(uni) == (27 if case[0] == 3 else 9)
    assert slots == -(-oc.n_free // 64) * 64 * npts
    K.to_stencil()
    for trial in range(3):
        x = rng.standard_normal(oc.n_free)
        if trial == 2:
            bad = rng.choice(oc.n_free, size=min(5, oc.n_free), replace=False)
            x[bad[::2]] = np.inf
            x[bad[1::2]] = np.nan
        dx, y0, y1 = ctx.from_numpy(x), pb.vector(), pb.vector()
        K.spmv(dx, y0)
        K.spmv_stencil(dx, y1)
        a, b = y0.numpy(), y1.numpy()
        nan = np.isnan(a)
        assert np.array_equal(nan, np.isnan(b)), trial
        assert np.array_equal(a[~nan].view(np.uint64), b[~nan].view(np.uint64)), trial
        if trial == 2:
            assert nan.any() and np.isinf(a).any()
        if trial < 2:
            assert same(b, oc.spmv(K.values(), x))
    # a refresh after new values
    K.set_values(rng.standard_normal(oc.nnz))
    K.to_stencil()
    x = rng.standard_normal(oc.n_free)
    dx, y0, y1 = ctx.from_numpy(x), pb.vector(), pb.vector()
    K.spmv(dx, y0)
    K.spmv_stencil(dx, y1)
    assert same(y0.numpy(), y1.numpy())


@pytest.mark.parametrize("n", [0, 1, 63, 256, 257, 5000, 300001])
def test_dot_norm_bitexact(ctx, n):
    say("dot n", n)
    rng = np.random.default_rng(n)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    dx, dy = ctx.from_numpy(x), ctx.from_numpy(y)
    assert ctx.dot(dx, dy) == O.dot(x, y, O.ORDER_GPU)
    assert ctx.norm2(dx) == np.sqrt(O.dot(x, x, O.ORDER_GPU))
    if n:
        assert abs(ctx.dot(dx, dy) - np.dot(x, y)) <= 1e-12 * np.abs(x * y).sum()


def test_blas1(ctx):
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal(1000), rng.standard_normal(1000)
    dx, dy, dz = ctx.from_numpy(x), ctx.from_numpy(y), ctx.empty(1000)
    ctx.axpy(0.3, dx, dy)
    assert same(dy.numpy(), y + 0.3 * x)
    ctx.axpby(0.7, -1.5, dx, dy)
    assert same(dy.numpy(), -1.5 * (y + 0.3 * x) + 0.7 * x)
    ctx.scal(2.5, dx)
    assert same(dx.numpy(), x * 2.5)
    ctx.mul(dx, dy, dz)
    assert same(dz.numpy(), dx.numpy() * dy.numpy())
    ctx.copy(dz, dx)
    assert same(dx.numpy(), dz.numpy())
    ctx.fill(dx, 4.25)
    assert np.all(dx.numpy() == 4.25)


def test_blas1_f32(ctx):
    """FP32 vectors: the elementwise operations bitwise numpy's float32
    arithmetic (each product and sum rounded once); dot and norm2 sum the
    exact products in FP64 (a fixed tree, so to 1e-12 of numpy's FP64 dot)."""
    rng = np.random.default_rng(4)
    n = 3001
    x = rng.standard_normal(n).astype(np.float32)
    y = rng.standard_normal(n).astype(np.float32)
    dx, dy, dz = ctx.from_numpy(x), ctx.from_numpy(y), ctx.empty(n, np.float32)
    f = np.float32
    d64 = np.dot(x.astype(np.float64), y.astype(np.float64))
    assert abs(ctx.dot(dx, dy) - d64) <= 1e-12 * np.abs(x.astype(np.float64) * y).sum()
    n64 = np.sqrt(np.dot(x.astype(np.float64), x.astype(np.float64)))
    assert abs(ctx.norm2(dx) - n64) <= 1e-13 * n64
    ctx.axpy(0.3, dx, dy)
    y1 = y + f(0.3) * x
    assert same(dy.numpy(), y1) and dy.numpy().dtype == np.float32
    ctx.axpby(0.7, -1.5, dx, dy)
    y2 = f(-1.5) * y1 + f(0.7) * x
    assert same(dy.numpy(), y2)
    ctx.scal(2.5, dx)
    x2 = x * f(2.5)
    assert same(dx.numpy(), x2)
    ctx.mul(dx, dy, dz)
    assert same(dz.numpy(), x2 * y2)
    ctx.copy(dz, dx)
    assert same(dx.numpy(), x2 * y2)
    with pytest.raises(TypeError):
        ctx.axpy(1.0, dx, ctx.from_numpy(np.zeros(n)))


FORMSETS = {
    "diffusion": [(O.DIFFUSION, 1.0)],
    "mass": [(O.MASS, 1.0)],
    "be": [(O.DIFFUSION, 1.0), (O.MASS | O.SCALED, 1.0 / 3e-3)],
}


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("fs", sorted(FORMSETS))
def test_assemble_matrix_bitexact(ctx, case, fs):
    pb, oc = build(ctx, case, kappa_seed=7)
    forms = FORMSETS[fs]
    K = pb.matrix()
    pb.assemble_matrix(K, forms)
    got = K.values()
    assert same(got, oc.assemble_matrix(forms))
    oc.set_order(O.ORDER_REF)
    ref = oc.assemble_matrix(forms)
    assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("case", CASES + [(3, (5, 4, 9), (1, 1, 1), 2, [0, 4], ["x", "1"], None)])
def test_assembly_schemes_bitwise(ctx, case):
    """The two matrix assembly schemes -- row gather (element pass, then one
    pass per row summing its elements' rows in colour order) and the
    colour-by-colour element scatter -- give the same bits for every form
    set and for a temperature-dependent Jacobian; a product with each
    matrix agrees bitwise too, so the padding slots hold the same zeros;
    the essential lift by row gather equals the colour-by-colour lift.
    The gather runs first, then the scatter, then the gather again on a
    problem whose value arrays were written by the other scheme."""
    import residuum_hip as rh
    dim, n, p, nq, tags, exprs, ext = case
    pb, oc = build(ctx, case, kappa_seed=17)
    assert pb.set_assembly(1) == 1
    x = ctx.from_numpy(np.random.default_rng(3).standard_normal(oc.n_free))
    got = {}
    K = pb.matrix()
    for mode in (1, 0, 1):
        pb.set_assembly(mode)
        for fs in sorted(FORMSETS):
            pb.assemble_matrix(K, FORMSETS[fs])
            y = pb.vector()
            K.spmv(x, y)
            got.setdefault(fs, []).append((bits(K.values()), bits(y.numpy())))
    for fs, runs in got.items():
        say("form set", fs)
        for v, y in runs[1:]:
            assert np.array_equal(v, runs[0][0]) and np.array_equal(y, runs[0][1])
    # the essential lift: by row gather (two launches) or colour by colour
    lifts = []
    for mode in (1, 0):
        pb.set_assembly(mode)
        F = ctx.from_numpy(np.random.default_rng(8).standard_normal(oc.n_free))
        pb.apply_essential(F, FORMSETS["be"])
        lifts.append(bits(F.numpy()))
    assert np.array_equal(lifts[0], lifts[1])
    if dim == 2:
        tensor = (1.0, 0.2, 0.2, 0.1)
        pj = rh.Problem(ctx, dim, n, p, nq, tags, exprs, ext=ext)
        pj.set_conductivity(1.0, k_expr="1.0 + 0.5*T", dk_expr="0.5", tensor=tensor)
        U = ctx.from_numpy(np.random.default_rng(2).standard_normal(oc.n_free))
        forms = [(O.DIFFUSION, 1.0), (O.TDIFFUSION, 1.0)]
        J = pj.matrix()
        vals = []
        for mode in (1, 0):
            pj.set_assembly(mode)
            pj.assemble_matrix(J, forms, U=U)
            vals.append(bits(J.values()))
        assert np.array_equal(vals[0], vals[1])


def test_be_run_assembly_schemes_bitwise(ctx):
    """Backward Euler through the fused K + M/dt and M U_prev assembly,
    captured as a graph: both assembly schemes give the oracle's solution
    bit for bit, and the same mixed-precision solution."""
    case = (3, (6, 5, 10), (1, 1, 1), 2, [0, 1], ["1", "0"], None)
    pb, oc = build(ctx, case, kappa_seed=20261202)
    out = oc.be_run(0.0, 0.005, 1e-3, cg_tol=1e-8)
    mixed = []
    for mode in (1, 0):
        pb.set_assembly(mode)
        U = pb.vector(0.0)
        steps, iters = pb.be_run(U, 0.0, 0.005, 1e-3, tol=1e-8, use_graph=True)
        say("assembly", mode, "iters", list(iters))
        assert steps == out["steps"] == 5 and list(iters) == list(out["iters"])
        assert np.array_equal(bits(U.numpy()), bits(out["U"]))
        Um = pb.vector(0.0)
        steps, iters = pb.be_run(Um, 0.0, 0.005, 1e-3, tol=1e-8, mixed=True, use_graph=True)
        mixed.append((list(iters), bits(Um.numpy())))
    assert mixed[0][0] == mixed[1][0] and np.array_equal(mixed[0][1], mixed[1][1])


def test_matrix_outliving_its_problem(ctx):
    """Destroying a problem orphans the matrices made on it (a C++ solver's
    copy may outlive the problem): a product with such a matrix fails with
    RSD_ERR_ARG and says why, and its release still succeeds."""
    import residuum_hip as rh
    L = rh.lib()
    pb = rh.Problem(ctx, 2, (4, 4), (1, 1), 2, [0], ["0"])
    K = pb.matrix()
    pb.assemble_matrix(K, [(rh.DIFFUSION, 1.0)])
    x, y = pb.vector(1.0), pb.vector()
    rh.check(L.rsd_spmv_f64(ctx.h, K.h, rh.dptr(x), rh.dptr(y)))
    pb._matrices.discard(K)  # the binding would release the matrix first
    pb.close()
    assert L.rsd_spmv_f64(ctx.h, K.h, rh.dptr(x), rh.dptr(y)) == -1  # RSD_ERR_ARG
    assert b"destroyed" in L.rsd_last_error_string()
    assert L.rsd_matrix_destroy(K.h) == 0
    K.h = None


CONTRACT_CASES = [(3, (7, 6, 5), (1, 1, 1), 2, [0, 1], ["1", "0"], None),
                  (2, (9, 7), (2, 2), 3, [0, 1, 2, 3], ["(1-x)*(1-y)"] * 4, None)]


@pytest.mark.parametrize("case", CONTRACT_CASES)
def test_contracted_arithmetic(ctx, case):
    """RSD_ARITH_CONTRACTED, the Backward Euler form pair in fused
    multiply-adds, against the oracle's reference-order arithmetic:
    - matrix entries within 16 ulps of their row's largest |entry|, by both
      assembly schemes, which agree with each other bit for bit;
    - the matrix-free operand product within 16 ulps of sum_j |A_ij x_j|;
    - Backward Euler (10 steps, CG to 1e-13, the fused mass vector) and the
      matrix-free CG (to 1e-13) within 1e-11 of the oracle's solutions;
    - back in RSD_ARITH_REFERENCE, the matrix is the oracle's bit for bit."""
    import residuum_hip as rh
    pb, oc = build(ctx, case, kappa_seed=29)
    eps = np.finfo(np.float64).eps
    dt = 2e-3
    forms = [(O.DIFFUSION, 1.0), (O.MASS | O.SCALED, 1.0 / dt)]
    ref = oc.assemble_matrix(forms)
    rows = np.repeat(np.arange(oc.n_free), np.diff(oc.rp))
    rowmax = np.zeros(oc.n_free)
    np.maximum.at(rowmax, rows, np.abs(ref))
    pb.set_arithmetic(rh.ARITH_CONTRACTED)
    K = pb.matrix()
    vals = []
    for mode in (1, 0):
        pb.set_assembly(mode)
        pb.assemble_matrix(K, forms)
        vals.append(K.values())
    assert np.array_equal(bits(vals[0]), bits(vals[1]))
    err = np.max(np.abs(vals[0] - ref) / rowmax[rows]) / eps
    say(f"matrix: max error {err:.2f} ulps of the row's largest entry")
    assert err <= 16
    x = np.random.default_rng(6).standard_normal(oc.n_free)
    y = pb.vector()
    pb.fem_apply(y, forms, ctx.from_numpy(x))
    yo = oc.assemble_vector(forms, operand=x, selector=O.SEL_OPERAND)
    scale = np.zeros(oc.n_free)
    np.add.at(scale, rows, np.abs(ref * x[oc.ci]))
    erry = np.max(np.abs(y.numpy() - yo) / scale) / eps
    say(f"operand product: max error {erry:.2f} ulps of sum |A_ij x_j|")
    assert erry <= 16
    out = oc.be_run(0.0, 10 * dt, dt, cg_tol=1e-13, cg_max=2000)
    pb.set_assembly(1)
    U = pb.vector(0.0)
    steps, _ = pb.be_run(U, 0.0, 10 * dt, dt, tol=1e-13, max_iters=2000)
    eU = np.max(np.abs(U.numpy() - out["U"])) / np.max(np.abs(out["U"]))
    say(f"BE 10 steps: steps {steps}, max rel diff {eU:.2e}")
    assert steps == out["steps"] == 10 and eU <= 1e-11
    b = np.random.default_rng(7).standard_normal(oc.n_free)
    dx = pb.vector(0.0)
    rep = pb.cg_fem(ctx.from_numpy(b), dx, forms, tol=1e-13, max_iters=2000)
    xo, ro = oc.cg_fem(forms, b, tol=1e-13, max_iters=2000)
    ex = np.max(np.abs(dx.numpy() - xo)) / np.max(np.abs(xo))
    say(f"matrix-free CG: iters {rep.iterations} / {ro.iterations}, max rel diff {ex:.2e}")
    assert rep.converged and ex <= 1e-11
    pb.set_arithmetic(rh.ARITH_REFERENCE)
    pb.assemble_matrix(K, forms)
    assert np.array_equal(bits(K.values()), bits(ref))


@pytest.mark.parametrize("case", CONTRACT_CASES)
def test_uniform_geometry_arithmetic(ctx, case):
    """RSD_ARITH_UNIFORM on a block mesh: the Backward Euler pair's element
    matrix as kappa_e Kh + coef Mh (Kh, Mh of element 0), against the
    oracle's per-element reference order:
    - matrix entries within 1e-12 of their row's largest |entry| (the
      per-element coordinate rounding is about 1e-14 relative);
    - the operand product within 1e-12 of sum_j |A_ij x_j|;
    - Backward Euler (10 steps, CG to 1e-13, the fused mass vector) and the
      matrix-free CG (to 1e-13) within 1e-10 of the oracle's solutions, and
      the mixed-precision Backward Euler loop (with the row pass's product)
      within 1e-7;
    - on a mesh with one node moved (not uniform) the mode falls back to the
      contracted arithmetic, bit for bit."""
    import residuum_hip as rh
    dt = 2e-3
    # a Backward Euler run as the first work in this mode captures the step's
    # assembly in a graph: everything the mode allocates is reserved before
    pb0, _ = build(ctx, case, kappa_seed=31)
    pb0.set_arithmetic(rh.ARITH_UNIFORM)
    for mixed in (False, True):
        U0 = pb0.vector(0.0)
        pb0.be_run(U0, 0.0, 2 * dt, dt, tol=1e-10, use_graph=True, mixed=mixed)
    pb0.close()
    pb, oc = build(ctx, case, kappa_seed=31)
    forms = [(O.DIFFUSION, 1.0), (O.MASS | O.SCALED, 1.0 / dt)]
    ref = oc.assemble_matrix(forms)
    rows = np.repeat(np.arange(oc.n_free), np.diff(oc.rp))
    rowmax = np.zeros(oc.n_free)
    np.maximum.at(rowmax, rows, np.abs(ref))
    pb.set_arithmetic(rh.ARITH_UNIFORM)
    assert pb.set_assembly(1) == 1
    K = pb.matrix()
    pb.assemble_matrix(K, forms)
    err = np.max(np.abs(K.values() - ref) / rowmax[rows])
    say(f"uniform matrix: max error {err:.2e} of the row's largest entry")
    assert err <= 1e-12
    x = np.random.default_rng(6).standard_normal(oc.n_free)
    y = pb.vector()
    pb.fem_apply(y, forms, ctx.from_numpy(x))
    yo = oc.assemble_vector(forms, operand=x, selector=O.SEL_OPERAND)
    scale = np.zeros(oc.n_free)
    np.add.at(scale, rows, np.abs(ref * x[oc.ci]))
    erry = np.max(np.abs(y.numpy() - yo) / scale)
    say(f"uniform operand product: max error {erry:.2e} of sum |A_ij x_j|")
    assert erry <= 1e-12
    out = oc.be_run(0.0, 10 * dt, dt, cg_tol=1e-13, cg_max=2000)
    U = pb.vector(0.0)
    steps, _ = pb.be_run(U, 0.0, 10 * dt, dt, tol=1e-13, max_iters=2000)
    eU = np.max(np.abs(U.numpy() - out["U"])) / np.max(np.abs(out["U"]))
    Um = pb.vector(0.0)
    pb.be_run(Um, 0.0, 10 * dt, dt, tol=1e-10, max_iters=2000, mixed=True)
    eM = np.max(np.abs(Um.numpy() - out["U"])) / np.max(np.abs(out["U"]))
    say(f"uniform BE 10 steps: FP64 max rel diff {eU:.2e}, mixed {eM:.2e}")
    assert steps == out["steps"] == 10 and eU <= 1e-10 and eM <= 1e-7
    b = np.random.default_rng(7).standard_normal(oc.n_free)
    dx = pb.vector(0.0)
    rep = pb.cg_fem(ctx.from_numpy(b), dx, forms, tol=1e-13, max_iters=2000)
    xo, ro = oc.cg_fem(forms, b, tol=1e-13, max_iters=2000)
    ex = np.max(np.abs(dx.numpy() - xo)) / np.max(np.abs(xo))
    say(f"uniform matrix-free CG: iters {rep.iterations} / {ro.iterations}, max rel diff {ex:.2e}")
    assert rep.converged and ex <= 1e-10
    # one interior-free node moved by a hundredth of an element: not uniform
    dim, n, p, nq, tags, exprs, ext = case
    xyz = oc.xyz.copy().reshape(-1, dim)
    node = oc.ien.reshape(oc.n_elem, -1)[oc.n_elem // 2][0]
    xyz[node, 0] += 0.01 / n[0]
    got = []
    for arith in (rh.ARITH_UNIFORM, rh.ARITH_CONTRACTED):
        q = rh.Problem.from_mesh(ctx, dim, p, nq, xyz.ravel(), oc.ien, oc.rng, tags, exprs)
        q.set_conductivity(1.0, kappa=oc.kappa)
        q.set_arithmetic(arith)
        Kq = q.matrix()
        q.assemble_matrix(Kq, forms)
        yq = q.vector()
        q.fem_apply(yq, forms, ctx.from_numpy(x))
        got.append((bits(Kq.values()), bits(yq.numpy())))
        Kq.close()
        q.close()
    assert np.array_equal(got[0][0], got[1][0]) and np.array_equal(got[0][1], got[1][1])


FACTORED_CASES = [(2, (9, 7), (1, 1), 2, [0, 1, 2, 3], ["x*y"] * 4, None),
                  (3, (7, 6, 5), (1, 1, 1), 2, [0, 1], ["1", "0"], None),
                  (2, (6, 5), (1, 1), 3, [0, 2], ["1", "x"], None)]


def distorted(ctx, case, seed=3, amp=0.25):
    """A block mesh with its interior nodes moved by up to amp of an element
    and the whole mesh sheared (x += 0.3 y), so that no element's Jacobian is
    symmetric or constant: the GPU problem and the oracle on it."""
    import residuum_hip as rh
    dim, n, p, nq, tags, exprs, ext = case
    oc = O.Case(dim, n, p, nq, tags, exprs, ext=ext, order=O.ORDER_GPU)
    rng = np.random.default_rng(seed)
    xyz = oc.xyz.reshape(-1, dim)
    inner = np.all((xyz > 1e-12) & (xyz < 1 - 1e-12), axis=1)
    xyz[inner] += rng.uniform(-amp, amp, (int(inner.sum()), dim)) / max(n)
    xyz[:, 0] += 0.3 * xyz[:, 1]
    oc.kappa = np.exp(0.5 * rng.standard_normal(oc.n_elem))
    oc.set_order(O.ORDER_GPU)
    pb = rh.Problem.from_mesh(ctx, dim, p, nq, oc.xyz, oc.ien, oc.rng, tags, exprs)
    pb.set_conductivity(1.0, kappa=oc.kappa)
    return pb, oc


@pytest.mark.parametrize("case", FACTORED_CASES)
def test_factored_arithmetic(ctx, case):
    """RSD_ARITH_FACTORED on distorted Q1 meshes: the Backward Euler pair's
    operand product from the elements' multilinear coefficients (no element
    matrix), against the oracle's reference order on the same mesh:
    - the product within 1e-12 of sum_j |A_ij x_j| (a CPU restatement of the
      same arithmetic measured within 1.3e-15), for the operand, value and
      rate selectors' operand;
    - the matrix-free CG (to 1e-13) within 1e-10 of the oracle's solution;
    - the matrix and diagonal are the contracted arithmetic's, bit for bit;
    - on Q2 quads the product falls back to the contracted arithmetic, bit
      for bit."""
    import residuum_hip as rh
    dt = 2e-3
    forms = [(O.DIFFUSION, 1.0), (O.MASS | O.SCALED, 1.0 / dt)]
    say("case", case)
    pb, oc = distorted(ctx, case)
    ref = oc.assemble_matrix(forms)
    rows = np.repeat(np.arange(oc.n_free), np.diff(oc.rp))
    x = np.random.default_rng(6).standard_normal(oc.n_free)
    yo = oc.assemble_vector(forms, operand=x, selector=O.SEL_OPERAND)
    scale = np.zeros(oc.n_free)
    np.add.at(scale, rows, np.abs(ref * x[oc.ci]))
    pb.set_arithmetic(rh.ARITH_FACTORED)
    assert pb.set_assembly(1) == 1
    y = pb.vector()
    pb.fem_apply(y, forms, ctx.from_numpy(x))
    err = np.max(np.abs(y.numpy() - yo) / scale)
    say(f"factored operand product: max error {err:.2e} of sum |A_ij x_j|")
    assert err <= 1e-12
    b = np.random.default_rng(7).standard_normal(oc.n_free)
    dx = pb.vector(0.0)
    rep = pb.cg_fem(ctx.from_numpy(b), dx, forms, tol=1e-13, max_iters=4000)
    xo, ro = oc.cg_fem(forms, b, tol=1e-13, max_iters=4000)
    ex = np.max(np.abs(dx.numpy() - xo)) / np.max(np.abs(xo))
    say(f"factored matrix-free CG: iters {rep.iterations} / {ro.iterations}, max rel diff {ex:.2e}")
    assert rep.converged and ex <= 1e-10
    K = pb.matrix()
    pb.assemble_matrix(K, forms)
    kf = K.values()
    pb.set_arithmetic(rh.ARITH_CONTRACTED)
    pb.assemble_matrix(K, forms)
    assert np.array_equal(bits(kf), bits(K.values()))
    K.close()
    pb.close()
    # Q2 quads: the contracted product
    dim, n, p, nq, tags, exprs, ext = case
    if dim != 2:
        return
    pq, oq = distorted(ctx, (dim, n, (2, 2), 3, tags, exprs, ext), seed=5, amp=0.1)
    xq = np.random.default_rng(8).standard_normal(oq.n_free)
    got = []
    for arith in (rh.ARITH_FACTORED, rh.ARITH_CONTRACTED):
        pq.set_arithmetic(arith)
        yq = pq.vector()
        pq.fem_apply(yq, forms, ctx.from_numpy(xq))
        got.append(bits(yq.numpy()))
    assert np.array_equal(got[0], got[1])
    pq.close()


@pytest.mark.parametrize("case", CASES[:2])
def test_assemble_jacobian_tdep_bitexact(ctx, case):
    import residuum_hip as rh
    dim, n, p, nq, tags, exprs, ext = case
    tensor = (1.0, 0.2, 0.2, 0.1) if dim == 2 else (1, 0.2, 0, 0.2, 0.1, 0, 0, 0, 1)
    pb = rh.Problem(ctx, dim, n, p, nq, tags, exprs, ext=ext)
    pb.set_conductivity(1.0, k_expr="1.0 + 0.5*T", dk_expr="0.5", tensor=tensor)
    oc = O.Case(dim, n, p, nq, tags, exprs, ext=ext, order=O.ORDER_GPU, k_expr="1.0 + 0.5*T", dk_expr="0.5",
                tensor=tensor)
    U = np.random.default_rng(2).standard_normal(oc.n_free)
    forms = [(O.DIFFUSION, 1.0), (O.TDIFFUSION, 1.0)]
    K = pb.matrix()
    pb.assemble_matrix(K, forms, U=ctx.from_numpy(U))
    assert same(K.values(), oc.assemble_matrix(forms, U=U))
    R = pb.vector()
    pb.assemble_vector(R, [(O.DIFFUSION, 1.0)], U=ctx.from_numpy(U))
    assert same(R.numpy(), oc.assemble_vector([(O.DIFFUSION, 1.0)], U=U))


@pytest.mark.parametrize("case", CASES)
def test_assemble_vectors_bitexact(ctx, case):
    pb, oc = build(ctx, case, kappa_seed=11)
    rng = np.random.default_rng(5)
    U = rng.standard_normal(oc.n_free)
    dU = ctx.from_numpy(U)
    F = pb.vector()
    for forms in ([(O.DIFFUSION, 1.0)], [(O.MASS, 1.0)], [(O.DIFFUSION, 1.0), (O.MASS | O.SCALED, 250.0)]):
        say("vector forms", forms)
        pb.assemble_vector(F, forms, U=dU)
        assert same(F.numpy(), oc.assemble_vector(forms, U=U))
        pb.assemble_vector(F, forms, U=dU, operand=dU, selector=O.SEL_OPERAND)
        assert same(F.numpy(), oc.assemble_vector(forms, U=U, operand=U, selector=O.SEL_OPERAND))
    src = "x*y*z" if case[0] == 3 else "x*y"
    pb.assemble_load(F, src=src)
    assert same(F.numpy(), oc.load(src))
    nod = rng.random(oc.n_nodes)
    pb.assemble_load(F, nodal=ctx.from_numpy(nod))
    assert same(F.numpy(), oc.load(nodal=nod))
    pb.assemble_load(F, src="10*sin(pi*x)*sin(pi*y)")
    ref = oc.load("10*sin(pi*x)*sin(pi*y)")
    assert np.max(np.abs(F.numpy() - ref)) <= 1e-14 * max(1.0, np.max(np.abs(ref)))
    for forms in ([(O.DIFFUSION, 1.0)], [(O.DIFFUSION, 1.0), (O.MASS | O.SCALED, 250.0)]):
        F.upload(U)
        pb.apply_essential(F, forms)
        assert same(F.numpy(), oc.lift(forms, U))


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("precond", [0, 1])
def test_cg_bitexact(ctx, case, precond):
    pb, oc = build(ctx, case, kappa_seed=3)
    forms = [(O.DIFFUSION, 1.0), (O.MASS | O.SCALED, 100.0)]
    val = oc.assemble_matrix(forms)
    K = pb.matrix()
    pb.assemble_matrix(K, forms)
    rng = np.random.default_rng(9)
    b = rng.standard_normal(oc.n_free)
    x0 = rng.standard_normal(oc.n_free) * 0.1
    db, dx = ctx.from_numpy(b), ctx.from_numpy(x0)
    rep, hist = K.cg(db, dx, tol=1e-10, max_iters=500, precond=precond, hist=True)
    xo, ro, ho = oc.cg(val, b, x0=x0, tol=1e-10, max_iters=500, precond=precond, hist=True)
    say("iters", rep.iterations, ro.iterations)
    assert rep.converged == ro.converged == 1
    assert rep.iterations == ro.iterations
    assert same(dx.numpy(), xo)
    assert same(hist[:rep.iterations + 1], ho[:ro.iterations + 1])
    assert rep.final_residual == ro.final_residual and rep.initial_residual == ro.initial_residual
    # against the reference-order oracle (tightly converged)
    oc.set_order(O.ORDER_REF)
    xr, rr = oc.cg(oc.assemble_matrix(forms), b, x0=x0, tol=1e-13, max_iters=2000, precond=precond)
    dx.upload(x0)
    K.cg(db, dx, tol=1e-13, max_iters=2000, precond=precond)
    assert np.max(np.abs(dx.numpy() - xr)) <= 1e-12 * np.max(np.abs(xr))


@pytest.mark.parametrize("case", CASES)
def test_assemble_diagonal_bitexact(ctx, case):
    """assembleDiagonal (Assembler.tpp:142-197) on the device: bitwise the
    oracle in GPU order and bitwise the diagonal of the assembled matrix."""
    pb, oc = build(ctx, case, kappa_seed=6)
    rp, ci = pb.pattern()
    ds = np.array([rp[i] + int(np.searchsorted(ci[rp[i]:rp[i + 1]], i)) for i in range(oc.n_free)])
    K = pb.matrix()
    d = pb.vector()
    for forms in ([(O.DIFFUSION, 1.0)], [(O.MASS, 1.0)], [(O.DIFFUSION, 1.0), (O.MASS | O.SCALED, 250.0)]):
        pb.assemble_diagonal(d, forms)
        pb.assemble_matrix(K, forms)
        got = d.numpy()
        assert same(got, oc.assemble_diagonal(forms))
        assert same(got, K.values()[ds])


@pytest.mark.parametrize("case", [CASES[1], CASES[2], CASES[3]])
@pytest.mark.parametrize("precond", [0, 1])
def test_cg_fem_bitexact(ctx, case, precond):
    """CG over the matrix-free FEMOperator: apply by element assembly, Jacobi
    from assembleDiagonal.  Iterates, iteration count and residual history
    bitwise the oracle's FEM CG in GPU order; the solution equals the CSR
    solve's to the solver tolerance."""
    pb, oc = build(ctx, case, kappa_seed=8)
    forms = [(O.DIFFUSION, 1.0), (O.MASS | O.SCALED, 100.0)]
    rng = np.random.default_rng(10)
    b = rng.standard_normal(oc.n_free)
    x0 = rng.standard_normal(oc.n_free) * 0.1
    db, dx = ctx.from_numpy(b), ctx.from_numpy(x0)
    rep, hist = pb.cg_fem(db, dx, forms, tol=1e-10, max_iters=500, precond=precond, hist=True)
    xo, ro, ho = oc.cg_fem(forms, b, x0=x0, tol=1e-10, max_iters=500, precond=precond, hist=True)
    say("fem iters", rep.iterations, ro.iterations)
    assert rep.converged == ro.converged == 1 and rep.iterations == ro.iterations
    assert same(dx.numpy(), xo)
    assert same(hist[:rep.iterations + 1], ho[:ro.iterations + 1])
    K = pb.matrix()
    pb.assemble_matrix(K, forms)
    dx.upload(x0)
    K.cg(db, dx, tol=1e-13, max_iters=2000, precond=precond)
    xc = dx.numpy()
    dx.upload(x0)
    pb.cg_fem(db, dx, forms, tol=1e-13, max_iters=2000, precond=precond)
    assert np.max(np.abs(dx.numpy() - xc)) <= 1e-11 * np.max(np.abs(xc))


def test_matrix_free_cg_exact_bilinear(ctx):
    """The reference's matrix-free CG test on the device
    (tests/integration/full/heateq/cpu/HeatEquationMinimal2D.cpp:249-320):
    24x24 Q1 on [-1,1]^2, g = (1-x)(1-y), CG + Identity to 1e-12."""
    import residuum_hip as rh
    g = "(1-x)*(1-y)"
    case = (2, (24, 24), (1, 1), 2, [0, 1, 2, 3], [g] * 4, (-1.0, 1.0, -1.0, 1.0))
    pb, oc = build(ctx, case)
    assert pb.n_free == 529
    F = pb.vector()
    pb.assemble_load(F)
    pb.apply_essential(F, [(O.DIFFUSION, 1.0)])
    x = pb.vector(0.0)
    rep = pb.cg_fem(F, x, [(O.DIFFUSION, 1.0)], tol=1e-12, max_iters=10000, precond=rh.PC_IDENTITY)
    assert rep.converged == 1 and abs(rep.final_residual - 1e-12) <= 1e-11
    xy = oc.xyz.reshape(-1, 2)[oc.a2t]
    assert np.max(np.abs(x.numpy() - (1 - xy[:, 0]) * (1 - xy[:, 1]))) <= 1e-10


@pytest.mark.parametrize("case", CASES)
def test_boundary_reaction_bitexact(ctx, case):
    """Weak boundary reaction (QuantityEvaluator.tpp:95-199, the heat pairs)
    on the device: result and face measure per tag bitwise the oracle's
    serial sums, steady and transient, with and without a source; the tag
    list holds every face tag, a repeated tag (the last entry takes it) and
    one no face carries."""
    pb, oc = build(ctx, case, kappa_seed=9)
    dim = case[0]
    tags = list(range(2 * dim)) + [1, 17]
    rng = np.random.default_rng(12)
    U, Ud = rng.standard_normal(oc.n_free), rng.standard_normal(oc.n_free)
    dU, dUd = ctx.from_numpy(U), ctx.from_numpy(Ud)
    for src in (None, "x*y + 2*t"):
        for rate in (False, True):
            got = pb.boundary_reaction(tags, dU, Udot=dUd if rate else None, src=src, time=0.5)
            want = oc.boundary_reaction(tags, U, Udot=Ud if rate else None, src=src, time=0.5)
            say("reaction", src, rate, got[0][:3], want[0][:3])
            assert same(got[0], want[0]) and same(got[1], want[1])
    assert got[0][1] == 0.0 and got[1][1] == 0.0 and got[1][-1] == 0.0


@pytest.mark.parametrize("case", CASES)
def test_boundary_flux_bitexact(ctx, case):
    """Strong boundary flux (QuantityEvaluator.tpp:14-93, HeatFluxIntegrand)
    on the device: integral and face measure per tag bitwise the oracle's
    serial sums (per-element conductivity, Dirichlet values at `time`)."""
    pb, oc = build(ctx, case, kappa_seed=4)
    tags = list(range(2 * case[0])) + [1, 17]
    U = np.random.default_rng(14).standard_normal(oc.n_free)
    got = pb.boundary_flux(tags, ctx.from_numpy(U), time=0.25)
    want = oc.boundary_flux(tags, U, time=0.25)
    say("flux", got[0][:3], want[0][:3])
    assert same(got[0], want[0]) and same(got[1], want[1])
    # the measure is the same sum as the weak reaction's
    assert same(got[1], pb.boundary_reaction(tags, ctx.from_numpy(U), time=0.25)[1])


NATURAL_ENTRIES = {2: [(1, 0), (2, 1), (3, "nodal"), (2, "nodal"), (0, 1)],
                   3: [(0, 1), (1, 0), (5, "nodal"), (2, 1), (3, "nodal"), (2, 0)]}


def _device_entries(ctx, oc, spec, seed):
    """(tag, flux) lists for the oracle and the device: kinds become the
    reference harness's flux expressions, "nodal" a seeded nodal field."""
    rng = np.random.default_rng(seed)
    ho, dv = [], []
    for tag, fl in spec:
        if fl == "nodal":
            a = rng.standard_normal(oc.n_nodes * oc.dim)
            ho.append((tag, a))
            dv.append((tag, ctx.from_numpy(a)))
        else:
            ho.append((tag, fl))
            dv.append((tag, O.FLUX_KINDS[fl][:oc.dim]))
    return ho, dv


@pytest.mark.parametrize("case", CASES)
def test_natural_bc_bitexact(ctx, case):
    """Natural (flux) boundary conditions (BoundaryApplicator.tpp:66-140,
    FluxBoundaryForm / NodalFluxForm) on the device: F after the face terms
    bitwise the oracle's serial loop (pinned against the reference's own
    applyNaturalBCs), on constrained and free faces, several entries on one
    tag, at two times; the sin/cos flux within 1e-14 (device libm)."""
    pb, oc = build(ctx, case)
    ho, dv = _device_entries(ctx, oc, NATURAL_ENTRIES[case[0]], 21)
    F0 = np.random.default_rng(22).standard_normal(oc.n_free)
    for time in (0.0, 0.4):
        F = ctx.from_numpy(F0)
        pb.apply_natural(F, dv, time=time)
        want = oc.apply_natural(ho, F0, time=time)
        say("natural", time, F.numpy()[:3], want[:3])
        assert same(F.numpy(), want)
    # the time-dependent, transcendental flux (kind 2)
    F = ctx.from_numpy(F0)
    pb.apply_natural(F, [(1, O.FLUX_KINDS[2][:oc.dim]), (3, O.FLUX_KINDS[2][:oc.dim])], time=0.3)
    want = oc.apply_natural([(1, 2), (3, 2)], F0, time=0.3)
    assert np.max(np.abs(F.numpy() - want)) <= 1e-14 * np.max(np.abs(want))
    # no entries: F unchanged; a flux with the wrong arity is refused
    F = ctx.from_numpy(F0)
    pb.apply_natural(F, [])
    assert same(F.numpy(), F0)
    import residuum_hip as rh
    with pytest.raises(rh.RsdError):
        pb.apply_natural(F, [(0, ("1",) * (oc.dim + 1))])


def test_boundary_flux_nonlinear_anisotropic_bitexact(ctx):
    """k(T) times an anisotropic tensor: K = k(T) shape at each face point."""
    import residuum_hip as rh
    dim, n, p, nq = 2, (5, 6), (2, 2), 3
    tensor = (2.0, 0.5, 0.5, 1.0)
    pb = rh.Problem(ctx, dim, n, p, nq, [1, 2], ["x*y", "1 - x"])
    pb.set_conductivity(1.0, k_expr="1.0 + 0.5*T", dk_expr="0.5", tensor=tensor)
    oc = O.Case(dim, n, p, nq, [1, 2], ["x*y", "1 - x"], order=O.ORDER_GPU, k_expr="1.0 + 0.5*T", dk_expr="0.5",
                tensor=tensor)
    U = np.random.default_rng(15).standard_normal(oc.n_free)
    got = pb.boundary_flux([0, 1, 2, 3], ctx.from_numpy(U), time=1.0)
    want = oc.boundary_flux([0, 1, 2, 3], U, time=1.0)
    assert same(got[0], want[0]) and same(got[1], want[1])


def test_boundary_flux_reference_fixture(ctx):
    """tests/unit/equation/heateq/QuantityEvaluator.cpp:101-238 on the
    device: 4x4 Q1 on [-1,1]^2, k = 2, T = x + 2y; the left side's average
    k a, integral and length; the four sides' integrals and combinations;
    the same with T as a Dirichlet value on every side."""
    import residuum_hip as rh
    for tags in ([], [0, 1, 2, 3]):
        pb = rh.Problem(ctx, 2, (4, 4), (1, 1), 2, tags, ["x + 2*y"] * len(tags), ext=(-1.0, 1.0, -1.0, 1.0))
        pb.set_conductivity(2.0)
        oc = O.Case(2, (4, 4), (1, 1), 2, tags, ["x + 2*y"] * len(tags), ext=(-1.0, 1.0, -1.0, 1.0))
        xy = oc.xyz.reshape(-1, 2)[oc.a2t]
        dU = ctx.from_numpy(xy[:, 0] + 2 * xy[:, 1]) if oc.n_free else ctx.zeros(1)
        res, meas = pb.boundary_flux([0, 1, 2, 3], dU)
        assert np.max(np.abs(res - np.array([4.0, -4.0, 8.0, -8.0]))) <= 1e-10
        assert np.max(np.abs(meas - 2.0)) <= 1e-10
        assert abs(res[0] / meas[0] - 2.0) <= 1e-10
        assert abs(res[1] + res[2] - 4.0) <= 1e-10 and abs(res[1] - res[3] - 4.0) <= 1e-10
        assert abs(res.sum()) <= 1e-9


def test_boundary_reaction_nonlinear_bitexact(ctx):
    """k(T) conductivity: the stiffness pair uses k at each point's T."""
    import residuum_hip as rh
    dim, n, p, nq = 2, (6, 5), (2, 2), 3
    pb = rh.Problem(ctx, dim, n, p, nq, [0, 3], ["x", "1"])
    pb.set_conductivity(1.0, k_expr="1.0 + 0.5*T", dk_expr="0.5")
    oc = O.Case(dim, n, p, nq, [0, 3], ["x", "1"], order=O.ORDER_GPU, k_expr="1.0 + 0.5*T", dk_expr="0.5")
    U = np.random.default_rng(13).standard_normal(oc.n_free)
    got = pb.boundary_reaction([0, 1, 2, 3], ctx.from_numpy(U), src="3*x")
    want = oc.boundary_reaction([0, 1, 2, 3], U, src="3*x")
    assert same(got[0], want[0]) and same(got[1], want[1])


def test_boundary_reaction_heat_flux(ctx):
    """The reference's heat-flux fixture (tests/unit/equation/heateq/
    QuantityEvaluator.cpp:20-100) on the device: k = 2, T = x + 2y, no
    source; the reaction on x = -1 is k dT/dx times the edge length and the
    measure is the edge length."""
    import residuum_hip as rh
    pb = rh.Problem(ctx, 2, (8, 8), (1, 1), 2, [], [], ext=(-1.0, 1.0, -1.0, 1.0))
    pb.set_conductivity(2.0)
    oc = O.Case(2, (8, 8), (1, 1), 2, [], [], ext=(-1.0, 1.0, -1.0, 1.0))
    xy = oc.xyz.reshape(-1, 2)[oc.a2t]
    res, meas = pb.boundary_reaction([0, 2], ctx.from_numpy(xy[:, 0] + 2 * xy[:, 1]))
    assert abs(meas[0] - 2.0) < 1e-14 and abs(meas[1] - 2.0) < 1e-14
    assert abs(res[0] - 2.0 * 1.0 * 2.0) < 1e-12 and abs(res[1] - 2.0 * 2.0 * 2.0) < 1e-12


def test_cg_absolute_tolerance(ctx):
    """cg/Solver.tpp:45-48 with ToleranceType::Absolute: stop at the first k
    with ||r_k|| < tol, leaving exactly the iterate of k steps."""
    import residuum_hip as rh
    pb, oc = build(ctx, CASES[2], kappa_seed=5)
    forms = [(O.DIFFUSION, 1.0), (O.MASS | O.SCALED, 10.0)]
    K = pb.matrix()
    pb.assemble_matrix(K, forms)
    b = np.random.default_rng(4).standard_normal(oc.n_free)
    db, dx = ctx.from_numpy(b), ctx.empty(oc.n_free)
    ctx.fill(dx, 0.0)
    rep, hist = K.cg(db, dx, tol=1e-300, max_iters=60, hist=True)
    assert not rep.converged
    tau = 0.5 * (hist[12] + hist[13])
    want = next(k for k in range(len(hist)) if hist[k] < tau)
    ctx.fill(dx, 0.0)
    rep_a = K.cg(db, dx, tol=tau, max_iters=60, tol_type=rh.TOL_ABSOLUTE)
    xa = dx.numpy()
    assert rep_a.converged and rep_a.iterations == want and rep_a.final_residual < tau
    ctx.fill(dx, 0.0)
    K.cg(db, dx, tol=1e-300, max_iters=want + 1)  # stops on the iteration cap after `want` steps
    assert same(dx.numpy(), xa)
    # GMRES takes the same switch
    ctx.fill(dx, 0.0)
    rg = K.gmres(db, dx, tol=1e-6 * hist[0], max_iters=400, krylov=20, tol_type=rh.TOL_ABSOLUTE)
    assert rg.converged and rg.final_residual < 1e-6 * hist[0]


def test_cg_maxiter_and_zero_rhs(ctx):
    pb, oc = build(ctx, CASES[2])
    K = pb.matrix()
    pb.assemble_matrix(K, [(O.DIFFUSION, 1.0)])
    val = oc.assemble_matrix([(O.DIFFUSION, 1.0)])
    b = np.ones(oc.n_free)
    dx = pb.vector(0.0)
    rep = K.cg(ctx.from_numpy(b), dx, tol=1e-14, max_iters=3)
    xo, ro = oc.cg(val, b, tol=1e-14, max_iters=3)
    assert rep.converged == 0 and rep.iterations == 3 == ro.iterations
    assert same(dx.numpy(), xo) and rep.final_residual == ro.final_residual
    dx = pb.vector(0.0)
    rep = K.cg(pb.vector(0.0), dx, tol=1e-8)
    assert rep.converged == 1 and rep.iterations == 0


@pytest.mark.parametrize("use_graph", [False, True])
def test_be_run_bitexact(ctx, use_graph):
    case = (3, (8, 7, 6), (1, 1, 1), 2, [0, 1], ["1", "0"], None)
    pb, oc = build(ctx, case, kappa_seed=20261201)
    U = pb.vector(0.0)
    steps, iters = pb.be_run(U, 0.0, 0.01, 1e-3, tol=1e-8, use_graph=use_graph)
    out = oc.be_run(0.0, 0.01, 1e-3, cg_tol=1e-8)
    say("steps", steps, "iters", list(iters), list(out["iters"]))
    assert steps == out["steps"] == 10
    assert list(iters) == list(out["iters"])
    assert same(U.numpy(), out["U"])


@pytest.mark.parametrize("case", CASES[:3])
@pytest.mark.parametrize("precond", [0, 1])
def test_gmres_bitexact(ctx, case, precond):
    dim, n, p, nq, tags, exprs, ext = case
    import residuum_hip as rh
    tensor = (1.0, 0.2, 0.2, 0.1) if dim == 2 else (1, 0.2, 0, 0.2, 0.1, 0, 0, 0, 1)
    pb = rh.Problem(ctx, dim, n, p, nq, tags, exprs, ext=ext)
    pb.set_conductivity(1.0, k_expr="1.0 + 0.5*T", dk_expr="0.5", tensor=tensor)
    oc = O.Case(dim, n, p, nq, tags, exprs, ext=ext, order=O.ORDER_GPU, k_expr="1.0 + 0.5*T", dk_expr="0.5",
                tensor=tensor)
    U = np.random.default_rng(4).random(oc.n_free)
    forms = [(O.DIFFUSION, 1.0), (O.TDIFFUSION, 1.0)]
    J = pb.matrix()
    pb.assemble_matrix(J, forms, U=ctx.from_numpy(U))
    val = oc.assemble_matrix(forms, U=U)
    b = np.random.default_rng(5).standard_normal(oc.n_free)
    # the products read the slice-union copy on all three meshes
    has_su, su_slots = pb.layout_union()
    assert has_su == 1 and su_slots >= pb.nnz
    for krylov in (5, 30):
        dx = pb.vector(0.0)
        rep = J.gmres(ctx.from_numpy(b), dx, tol=1e-10, max_iters=2000, krylov=krylov, precond=precond)
        xo, ro = oc.gmres(val, b, tol=1e-10, max_iters=2000, krylov=krylov, precond=precond)
        say("gmres", krylov, rep.iterations, ro.iterations)
        assert rep.converged == ro.converged == 1 and rep.iterations == ro.iterations
        assert rep.final_residual == ro.final_residual
        assert same(dx.numpy(), xo)


@pytest.mark.parametrize("case", CASES[:3])
@pytest.mark.parametrize("precond", [0, 1])
def test_gmres_cgs2(ctx, case, precond):
    """RSD_ORTHO_CGS2 (classical Gram-Schmidt, reorthogonalised once) on the
    non-symmetric Jacobians of test_gmres_bitexact: not bitwise the
    reference's modified Gram-Schmidt, so against a dense solve -- converged
    at 1e-12 with the recomputed residual below 1e-11 of ||b||, the solution
    within 1e-9 of the dense one, and iteration counts within 5 % (at least
    2) of the oracle's MGS.  Krylov dimensions past 32 take the multi-vector
    passes in two chunks."""
    dim, n, p, nq, tags, exprs, ext = case
    import residuum_hip as rh
    tensor = (1.0, 0.2, 0.2, 0.1) if dim == 2 else (1, 0.2, 0, 0.2, 0.1, 0, 0, 0, 1)
    pb = rh.Problem(ctx, dim, n, p, nq, tags, exprs, ext=ext)
    pb.set_conductivity(1.0, k_expr="1.0 + 0.5*T", dk_expr="0.5", tensor=tensor)
    oc = O.Case(dim, n, p, nq, tags, exprs, ext=ext, order=O.ORDER_GPU, k_expr="1.0 + 0.5*T", dk_expr="0.5",
                tensor=tensor)
    U = np.random.default_rng(4).random(oc.n_free)
    forms = [(O.DIFFUSION, 1.0), (O.TDIFFUSION, 1.0)]
    J = pb.matrix()
    pb.assemble_matrix(J, forms, U=ctx.from_numpy(U))
    val = oc.assemble_matrix(forms, U=U)
    dense = np.zeros((oc.n_free, oc.n_free))
    rows = np.repeat(np.arange(oc.n_free), np.diff(oc.rp))
    dense[rows, oc.ci] = val
    b = np.random.default_rng(5).standard_normal(oc.n_free)
    xs = np.linalg.solve(dense, b)
    ctx.set_gmres_orthogonalization(rh.ORTHO_CGS2)
    try:
        for krylov in (5, 30, 40):
            dx = pb.vector(0.0)
            rep = J.gmres(ctx.from_numpy(b), dx, tol=1e-12, max_iters=4000, krylov=krylov, precond=precond)
            _, ro = oc.gmres(val, b, tol=1e-12, max_iters=4000, krylov=krylov, precond=precond)
            x = dx.numpy()
            err = np.max(np.abs(x - xs)) / np.max(np.abs(xs))
            res = np.linalg.norm(b - dense @ x) / np.linalg.norm(b)
            say("gmres cgs2", krylov, rep.iterations, ro.iterations, f"err {err:.2e} res {res:.2e}")
            assert rep.converged == 1
            assert abs(rep.iterations - ro.iterations) <= max(2, 0.05 * ro.iterations)
            assert res <= 1e-11 and err <= 1e-9
    finally:
        ctx.set_gmres_orthogonalization(rh.ORTHO_MGS)


@pytest.mark.parametrize("case", CASES[:3])
@pytest.mark.parametrize("precond", [0, 1])
@pytest.mark.parametrize("ortho", [0, 1], ids=["mgs", "cgs2"])
def test_gmres_mixed(ctx, case, precond, ortho):
    """Mixed-precision GMRES on the non-symmetric tangent Jacobian: its
    solution's FP64 relative residual, recomputed here from an FP64 product,
    is below the tolerance, which one FP32 inner solve cannot reach (so
    several refinements ran), and it agrees with the FP64 GMRES solution;
    with the inner GMRES orthogonalised by modified or by classical
    Gram-Schmidt reorthogonalised once (RSD_ORTHO_CGS2)."""
    import residuum_hip as rh
    dim, n, p, nq, tags, exprs, ext = case
    tensor = (1.0, 0.2, 0.2, 0.1) if dim == 2 else (1, 0.2, 0, 0.2, 0.1, 0, 0, 0, 1)
    pb = rh.Problem(ctx, dim, n, p, nq, tags, exprs, ext=ext)
    pb.set_conductivity(1.0, k_expr="1.0 + 0.5*T", dk_expr="0.5", tensor=tensor)
    nf = pb.n_free
    U = np.random.default_rng(4).random(nf)
    J = pb.matrix()
    pb.assemble_matrix(J, [(O.DIFFUSION, 1.0), (O.TDIFFUSION, 1.0)], U=ctx.from_numpy(U))
    b = np.random.default_rng(5).standard_normal(nf)
    db = ctx.from_numpy(b)
    ref = pb.vector(0.0)
    J.gmres(db, ref, tol=1e-13, max_iters=4000, krylov=30, precond=precond)
    for krylov in (5, 30, 40):
        dx = pb.vector(0.0)
        ctx.set_gmres_orthogonalization(ortho)
        try:
            rep, cyc = J.gmres_mixed(db, dx, tol=1e-10, max_iters=4000, krylov=krylov, precond=precond)
        finally:
            ctx.set_gmres_orthogonalization(rh.ORTHO_MGS)
        y = pb.vector()
        J.spmv(dx, y)
        rel = np.linalg.norm(b - y.numpy()) / np.linalg.norm(b)
        say("gmres mixed", krylov, precond, rep.iterations, cyc, rel)
        assert rep.converged == 1 and cyc >= 2 and rep.iterations > 0
        assert rep.final_residual_rel < 1e-10 and rel < 1.5e-10
        x = dx.numpy()
        assert np.max(np.abs(x - ref.numpy())) <= 1e-8 * np.max(np.abs(ref.numpy()))


def test_newton_mixed_inner(ctx):
    """Newton with mixed-precision inner solves (GMRES, and CG on the full
    FP32 copy of the non-symmetric Jacobian) on the anisotropic Q2 problem of
    the bitwise Newton test: the same Newton iteration count and the
    FP64-GMRES solution to 1e-9 (the inner solves stop at 1e-12 either way,
    so the iterates differ only by their rounding)."""
    import residuum_hip as rh
    dim, n, p, nq = 2, (12, 10), (2, 2), 3
    tensor = (1.0, 0.0, 0.0, 0.1)
    src = "10*sin(pi*x)*sin(pi*y)"
    out = []
    for lin_solver in (1, 2, 3):
        pb = rh.Problem(ctx, dim, n, p, nq, [0, 1, 2, 3], ["0"] * 4)
        pb.set_conductivity(1.0, k_expr="1.0 + 0.5*T", dk_expr="0.5", tensor=tensor)
        dU = pb.vector(0.0)
        its, res, lin = pb.newton(dU, src=src, abs_tol=1e-10, rel_tol=1e-10, lin_solver=lin_solver, lin_tol=1e-12,
                                  krylov=30)
        say("newton", lin_solver, its, list(lin), list(res))
        assert its > 0
        out.append((its, dU.numpy()))
    for its, u in out[1:]:
        assert its == out[0][0]
        assert np.max(np.abs(u - out[0][1])) <= 1e-9 * np.max(np.abs(out[0][1]))


def test_newton_bitexact(ctx):
    import residuum_hip as rh
    dim, n, p, nq = 2, (12, 10), (2, 2), 3
    tensor = (1.0, 0.0, 0.0, 0.1)
    pb = rh.Problem(ctx, dim, n, p, nq, [0, 1, 2, 3], ["0"] * 4)
    pb.set_conductivity(1.0, k_expr="1.0 + 0.5*T", dk_expr="0.5", tensor=tensor)
    oc = O.Case(dim, n, p, nq, [0, 1, 2, 3], ["0"] * 4, order=O.ORDER_GPU, k_expr="1.0 + 0.5*T", dk_expr="0.5",
                tensor=tensor)
    src = "10*sin(pi*x)*sin(pi*y)"
    dU = pb.vector(0.0)
    its, res, lin = pb.newton(dU, src=src, abs_tol=1e-10, rel_tol=1e-10, lin_solver=1, lin_tol=1e-12, krylov=30)
    Uo, its_o, res_o, lin_o = oc.newton(src=src, abs_tol=1e-10, rel_tol=1e-10, lin_solver=1, lin_tol=1e-12, krylov=30)
    say("newton", its, its_o, list(lin), list(lin_o))
    assert its == its_o and its > 0
    # the source uses sin: device and host libm may differ in the last bit
    assert np.max(np.abs(dU.numpy() - Uo)) <= 1e-12 * np.max(np.abs(Uo))
    assert np.max(np.abs(res - res_o)) <= 1e-12 * res_o[0]


def test_newton_kirchhoff_closed_form(ctx):
    import residuum_hip as rh
    pb = rh.Problem(ctx, 2, (40, 4), (1, 1), 2, [0, 1], ["0.0", "1.0"])
    pb.set_conductivity(1.0, k_expr="1.0 + 0.5*T", dk_expr="0.5")
    dU = pb.vector(0.0)
    its, res, lin = pb.newton(dU, abs_tol=1e-10, rel_tol=1e-10, lin_solver=0, lin_tol=1e-12)
    assert its > 0
    oc = O.Case(2, (40, 4), (1, 1), 2, [0, 1], ["0.0", "1.0"])
    x = oc.xyz.reshape(-1, 2)[oc.a2t][:, 0]
    theta = (-1.0 + np.sqrt(1.0 + 2 * 0.5 * 1.25 * x)) / 0.5
    assert np.max(np.abs(dU.numpy() - theta)) < 1e-6


@pytest.mark.parametrize("case", [CASES[2], CASES[3]])
def test_mixed_cg_reaches_fp64_residual(ctx, case):
    pb, oc = build(ctx, case, kappa_seed=17)
    forms = [(O.DIFFUSION, 1.0), (O.MASS | O.SCALED, 1.0 / 1e-3)]
    K = pb.matrix()
    pb.assemble_matrix(K, forms)
    val = oc.assemble_matrix(forms)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(oc.n_free)
    db = ctx.from_numpy(b)
    dx = pb.vector(0.0)
    rep, cycles = K.cg_mixed(db, dx, tol=1e-8)
    x = dx.numpy()
    r = b - oc.spmv(val, x)
    say("mixed", rep.iterations, cycles, rep.final_residual_rel)
    assert rep.converged == 1 and cycles >= 1
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-8
    xo, ro = oc.cg(val, b, tol=1e-12, max_iters=5000)
    assert np.max(np.abs(x - xo)) <= 1e-6 * np.max(np.abs(xo))
    # FP32 SpMV against the FP64 product
    K.to_f32()
    xv = rng.standard_normal(oc.n_free)
    d32, y32 = ctx.from_numpy(xv.astype(np.float32)), ctx.empty(oc.n_free, np.float32)
    K.spmv_f32(d32, y32)
    ref = oc.spmv(val.astype(np.float32).astype(np.float64), xv.astype(np.float32).astype(np.float64))
    assert np.max(np.abs(y32.numpy() - ref)) <= 1e-5 * np.max(np.abs(ref))


def permuted_problem(ctx, seed):
    """A 10x8x6 Q1 hex block mesh with randomly renumbered nodes: the rows
    no longer share at most 255 offset patterns, so every layout falls back
    to explicit columns."""
    import residuum_hip as rh
    xyz, ien, rng = O.block_mesh(3, (10, 8, 6), (0.0, 1.0, 0.0, 1.0, 0.0, 1.0), (1, 1, 1))
    nn = xyz.size // 3
    perm = np.random.default_rng(seed).permutation(nn)
    xyz2 = np.empty((nn, 3))
    xyz2[perm] = xyz.reshape(-1, 3)
    pb = rh.Problem.from_mesh(ctx, 3, (1, 1, 1), 2, xyz2.ravel(), perm[ien], rng, [0, 1], ["1", "0"])
    kap = np.exp(0.5 * O.normal_field(5, 0, pb.n_elem))
    pb.set_conductivity(1.0, kappa=kap)
    return pb


@pytest.mark.parametrize("precond", [0, 1])
def test_gmres_sliced_fallback_bitexact(ctx, precond):
    """GMRES on a renumbered mesh, whose rows share no offset patterns, so the
    products have no slice-union copy and walk the sliced layout's explicit
    columns: iterations, final residual and solution bitwise the oracle's
    GMRES in the GPU order on the same CSR matrix."""
    import ctypes as C
    pb = permuted_problem(ctx, 7)
    assert pb.layout_union() == (0, 0)
    forms = [(O.DIFFUSION, 1.0), (O.MASS | O.SCALED, 1.0 / 1e-3)]
    K = pb.matrix()
    pb.assemble_matrix(K, forms)
    rp, ci = pb.pattern()
    val = K.values()
    n = pb.n_free
    b = np.random.default_rng(8).standard_normal(n)
    for krylov in (5, 30):
        dx = pb.vector(0.0)
        rep = K.gmres(ctx.from_numpy(b), dx, tol=1e-10, max_iters=2000, krylov=krylov, precond=precond)
        xo = np.zeros(n)
        ro = O.OrcReport()
        O.lib().orc_gmres(C.c_int64(n), O.lp(rp), O.lp(ci), O.dp(val), O.dp(b), O.dp(xo), C.c_double(1e-10),
                          C.c_int64(2000), C.c_int64(krylov), precond, O.ORDER_GPU, C.byref(ro))
        say("gmres fallback", krylov, rep.iterations, ro.iterations)
        assert rep.converged == ro.converged == 1 and rep.iterations == ro.iterations
        assert rep.final_residual == ro.final_residual
        assert same(dx.numpy(), xo)


def symmetric_upper(A):
    """The operator of the FP32 half storage: every entry below the diagonal
    replaced by its mirror above it."""
    import scipy.sparse as sp
    return (sp.triu(A) + sp.triu(A, 1).T).tocsr()


@pytest.mark.parametrize("case", CASES + ["permuted"])
def test_fp32_half_storage(ctx, case):
    """The FP32 copy keeps each symmetric pair once when the rows share offset
    patterns: a product with a unit vector returns one column of the
    symmetrised operator, rounded to FP32, bit for bit (a single product per
    row, so no FP32 sum rounds).  A renumbered mesh keeps the full FP32 copy
    on explicit columns."""
    import scipy.sparse as sp
    forms = [(O.DIFFUSION, 1.0), (O.MASS | O.SCALED, 1.0 / 1e-3)]
    if case == "permuted":
        pb = permuted_problem(ctx, 4)
    else:
        pb, oc = build(ctx, case, kappa_seed=5)
    K = pb.matrix()
    pb.assemble_matrix(K, forms)
    rp, ci = pb.pattern()
    n = pb.n_free
    A = sp.csr_matrix((K.values(), ci, rp), shape=(n, n))
    if case != "permuted":
        assert np.array_equal(K.values(), oc.assemble_matrix(forms))
    mode32, npat, width, slots = pb.layout_f32()
    say("fp32 layout", mode32, npat, width, slots, "fp64 layout", pb.layout())
    if case == "permuted":
        assert pb.layout()[0] != 2 and mode32 == pb.layout()[0] and width == 0 and slots == pb.slots
        As = A
    else:
        upper = np.array([np.count_nonzero(ci[rp[i]:rp[i + 1]] >= i) for i in range(n)])
        if mode32 == 4:  # uniform stencil storage: one dense plane per non-negative offset
            offs = np.unique(ci - np.repeat(np.arange(n), np.diff(rp)))
            assert width == np.count_nonzero(offs >= 0) and len(offs) in (9, 27)
            assert slots >= ((n + 63) // 64) * 64 * width
        else:
            assert mode32 == 3 and width == upper.max()
            assert slots == ((n + 63) // 64) * 64 * width
        As = symmetric_upper(A)
    As32 = As.astype(np.float32)
    K.to_f32()
    y32 = ctx.empty(n, np.float32)
    for k in sorted({0, 1, n // 3, n // 2, n - 2, n - 1}):
        e = np.zeros(n, np.float32)
        e[k] = 1.0
        K.spmv_f32(ctx.from_numpy(e), y32)
        got = y32.numpy()
        want = np.asarray(As32[:, k].todense()).ravel()
        assert np.array_equal(got, want), (k, np.flatnonzero(got != want)[:8])
    xv = np.random.default_rng(11).standard_normal(n).astype(np.float32)
    K.spmv_f32(ctx.from_numpy(xv), y32)
    ref = As32.astype(np.float64) @ xv.astype(np.float64)
    assert np.max(np.abs(y32.numpy() - ref)) <= 1e-5 * np.max(np.abs(ref))
    # the mixed solve on either layout reaches the FP64 residual
    b = np.random.default_rng(12).standard_normal(n)
    x = pb.vector(0.0)
    rep, cycles = K.cg_mixed(ctx.from_numpy(b), x, tol=1e-8)
    r = b - A @ x.numpy()
    assert rep.converged == 1 and np.linalg.norm(r) / np.linalg.norm(b) < 1e-8


@pytest.mark.parametrize("case", CASES + ["permuted"])
def test_spmv_f32_padded_equals_per_pattern(ctx, case):
    """The mixed CG's own FP32 SpMV (uniform stencil: every row the same 27
    or 9 entries, zeros for the missing ones, x read three at a time from a
    zero-padded vector) is bitwise the per-pattern product, on every layout;
    the padding may hold only zeros and is never written."""
    forms = [(O.DIFFUSION, 1.0), (O.MASS | O.SCALED, 1.0 / 1e-3)]
    if case == "permuted":
        pb = permuted_problem(ctx, 6)
    else:
        pb, oc = build(ctx, case, kappa_seed=3)
    K = pb.matrix()
    pb.assemble_matrix(K, forms)
    K.to_f32()
    n = pb.n_free
    lo, hi = pb.padding_f32()
    mode32 = pb.layout_f32()[0]
    say("layout", mode32, "padding", lo, hi)
    assert (lo > 0 and hi > 0) == (mode32 == 4)
    rng = np.random.default_rng(16)
    for trial in range(3):
        xv = rng.standard_normal(n).astype(np.float32)
        xp = np.zeros(lo + n + hi, np.float32)
        xp[lo:lo + n] = xv
        dxp = ctx.from_numpy(xp)
        y1, y2 = ctx.empty(n, np.float32), ctx.empty(n, np.float32)
        K.spmv_f32(ctx.from_numpy(xv), y1)
        K.spmv_f32_padded(dxp, y2)
        assert np.array_equal(y1.numpy(), y2.numpy())
        assert np.array_equal(dxp.numpy(), xp)


def test_mixed_rejects_asymmetric_matrix(ctx):
    """A matrix whose mirrored pairs differ beyond rounding cannot go through
    the symmetric FP32 copy: the mixed solve fails with an argument error, and
    the same matrix made symmetric again solves."""
    import residuum_hip as rh
    pb, oc = build(ctx, CASES[2], kappa_seed=9)
    assert pb.layout_f32()[0] in (3, 4)
    forms = [(O.DIFFUSION, 1.0), (O.MASS | O.SCALED, 1.0 / 1e-3)]
    val = oc.assemble_matrix(forms)
    K = pb.matrix()
    i = oc.n_free // 2
    lo = oc.rp[i]  # the first entry of row i, left of the diagonal
    assert oc.ci[lo] < i
    bad = val.copy()
    bad[lo] *= 1.0 + 1e-6
    K.set_values(bad)
    b = ctx.from_numpy(np.random.default_rng(2).standard_normal(oc.n_free))
    x = pb.vector(0.0)
    with pytest.raises(rh.RsdError, match="not symmetric"):
        K.cg_mixed(b, x, tol=1e-8)
    K.set_values(val)
    x = pb.vector(0.0)
    rep, cycles = K.cg_mixed(b, x, tol=1e-8)
    assert rep.converged == 1


def test_be_run_mixed(ctx):
    case = (3, (10, 9, 8), (1, 1, 1), 2, [0, 1], ["1", "0"], None)
    pb, oc = build(ctx, case, kappa_seed=20261201)
    Um = pb.vector(0.0)
    steps_m, iters_m = pb.be_run(Um, 0.0, 0.01, 1e-3, tol=1e-8, mixed=True, use_graph=True)
    Ue = pb.vector(0.0)
    steps_e, _ = pb.be_run(Ue, 0.0, 0.01, 1e-3, tol=1e-8, mixed=True, use_graph=False)
    Ud = pb.vector(0.0)
    steps_d, iters_d = pb.be_run(Ud, 0.0, 0.01, 1e-3, tol=1e-8)
    say("mixed BE", list(iters_m), "fp64", list(iters_d))
    assert steps_m == steps_e == steps_d == 10
    assert same(Um.numpy(), Ue.numpy())
    ud = Ud.numpy()
    assert np.max(np.abs(Um.numpy() - ud)) <= 1e-6 * np.max(np.abs(ud))


def test_adjoint_bitexact_and_fd(ctx):
    import residuum_hip as rh
    case = (3, (5, 5, 4), (1, 1, 1), 2, [4], ["1"], None)
    dim, n, p, nq, tags, exprs, ext = case
    pb = rh.Problem(ctx, dim, n, p, nq, tags, exprs)
    oc = O.Case(dim, n, p, nq, tags, exprs, order=O.ORDER_GPU)
    theta = 0.5 * O.normal_field(20261203, 0, oc.n_elem)
    kappa = np.exp(theta)
    oc.kappa = kappa
    oc.set_order(O.ORDER_GPU)
    obs = np.arange(3, oc.n_free, 11)
    tobs = 0.2 + 0.01 * np.arange(obs.size) % 0.3
    dk, dU0, dg, dUN = ctx.from_numpy(kappa), pb.vector(0.0), ctx.empty(oc.n_elem), pb.vector()
    loss = pb.adjoint(dk, dU0, 0.0, 5e-3, 5, obs, tobs, tol=1e-13, d_grad=dg, d_UN=dUN)
    lo, go, UNo = oc.adjoint(0.0, 5e-3, 5, obs, tobs, cg_tol=1e-13)
    say("adjoint loss", loss, lo)
    assert loss == lo
    assert same(dUN.numpy(), UNo)
    assert same(dg.numpy(), go)
    # central differences along random directions (h = 1e-6)
    g = dg.numpy()
    rng = np.random.default_rng(20261206)
    for _ in range(4):
        d = rng.standard_normal(oc.n_elem)
        lp = pb.adjoint(ctx.from_numpy(np.exp(theta + 1e-6 * d)), dU0, 0.0, 5e-3, 5, obs, tobs, tol=1e-13,
                        want_grad=False)
        lm = pb.adjoint(ctx.from_numpy(np.exp(theta - 1e-6 * d)), dU0, 0.0, 5e-3, 5, obs, tobs, tol=1e-13,
                        want_grad=False)
        fd = (lp - lm) / 2e-6
        say("fd", fd, g @ d)
        assert abs(fd - g @ d) <= 1e-6 * abs(fd)


def test_adjoint_scaled_conductivity_repeated_observation(ctx):
    """k = 2 (the forward's element matrices are kappa_e * 2 * Khat) and an
    observation index listed twice: bitwise the oracle's adjoint, whose
    gradient is pinned by central differences on the CPU, and central
    differences of the device loss here."""
    import residuum_hip as rh
    dim, n, p, nq, tags, exprs = 3, (4, 5, 3), (1, 1, 1), 2, [4], ["1"]
    pb = rh.Problem(ctx, dim, n, p, nq, tags, exprs)
    oc = O.Case(dim, n, p, nq, tags, exprs, order=O.ORDER_GPU)
    theta = 0.4 * O.normal_field(11, 0, oc.n_elem)
    pb.set_conductivity(2.0)
    oc.k_scalar = 2.0
    oc.kappa = np.exp(theta)
    oc.set_order(O.ORDER_GPU)
    obs = np.concatenate([np.arange(1, oc.n_free, 5), [3, 3]]).astype(np.int64)
    tobs = np.linspace(0.05, 0.4, obs.size)
    dk, dU0, dg = ctx.from_numpy(np.exp(theta)), pb.vector(0.0), ctx.empty(oc.n_elem)
    loss = pb.adjoint(dk, dU0, 0.0, 5e-3, 3, obs, tobs, tol=1e-13, d_grad=dg)
    lo, go, _ = oc.adjoint(0.0, 5e-3, 3, obs, tobs, cg_tol=1e-13)
    assert loss == lo
    assert same(dg.numpy(), go)
    g = dg.numpy()
    rng = np.random.default_rng(5)
    for _ in range(2):
        d = rng.standard_normal(oc.n_elem)
        lp = pb.adjoint(ctx.from_numpy(np.exp(theta + 1e-6 * d)), dU0, 0.0, 5e-3, 3, obs, tobs, tol=1e-13,
                        want_grad=False)
        lm = pb.adjoint(ctx.from_numpy(np.exp(theta - 1e-6 * d)), dU0, 0.0, 5e-3, 3, obs, tobs, tol=1e-13,
                        want_grad=False)
        fd = (lp - lm) / 2e-6
        assert abs(fd - g @ d) <= 1e-6 * abs(fd)
    # a temperature-dependent conductivity is refused (the gradient is for a linear run)
    pb.set_conductivity(1.0, k_expr="1+0.5*T", dk_expr="0.5")
    with pytest.raises(rh.RsdError):
        pb.adjoint(dk, dU0, 0.0, 5e-3, 1, obs, tobs, tol=1e-13, d_grad=dg)


def test_adjoint_torch_autograd(ctx):
    torch = pytest.importorskip("torch")
    import residuum_hip as rh
    from residuum_hip.autograd import LogConductivityLoss
    pb = rh.Problem(ctx, 3, (4, 4, 4), (1, 1, 1), 2, [4], ["1"])
    theta = torch.tensor(0.3 * O.normal_field(5, 0, pb.n_elem), dtype=torch.float64, requires_grad=True)
    obs = np.arange(0, pb.n_free, 5)
    tobs = np.full(obs.size, 0.3)
    U0 = pb.vector(0.0)
    loss = LogConductivityLoss.apply(theta, pb, U0, 0.0, 5e-3, 3, obs, tobs, 1e-13)
    loss.backward()
    g = theta.grad.detach().numpy()
    oc = O.Case(3, (4, 4, 4), (1, 1, 1), 2, [4], ["1"], order=O.ORDER_GPU)
    oc.kappa = np.exp(theta.detach().numpy())
    oc.set_order(O.ORDER_GPU)
    lo, go, _ = oc.adjoint(0.0, 5e-3, 3, obs, tobs, cg_tol=1e-13)
    assert loss.item() == lo
    assert np.array_equal(g, go)


@pytest.mark.parametrize("use_graph", [False, True])
def test_batched_cg_equals_standalone(ctx, use_graph):
    import residuum_hip as rh
    sides = [3, 7, 5, 9, 4, 6]
    pbs, mats, bs = [], [], []
    for i, s in enumerate(sides):
        pb = rh.Problem(ctx, 3, (s, s, s), (1, 1, 1), 2, list(range(6)), ["0"] * 6)
        kap = np.exp(rh.philox_normal(20261202, i, pb.n_elem))
        pb.set_conductivity(1.0, kappa=kap)
        K = pb.matrix()
        pb.assemble_matrix(K, [(O.DIFFUSION, 1.0)])
        F = pb.vector()
        nod = ctx.from_numpy(rh.philox_uniform(20261202, 1000 + i, pb.n_nodes))
        pb.assemble_load(F, nodal=nod)
        pbs.append(pb); mats.append(K); bs.append(F.numpy())
    batch = rh.Batch(ctx, mats)
    db = ctx.from_numpy(batch.pack(bs))
    dx = ctx.from_numpy(np.zeros(batch.total_rows))
    reps = batch.cg(db, dx, tol=1e-8, max_iters=1000, use_graph=use_graph)
    xs = batch.unpack(dx.numpy())
    for i, (pb, K, b) in enumerate(zip(pbs, mats, bs)):
        x1 = pb.vector(0.0)
        r1 = K.cg(ctx.from_numpy(b), x1, tol=1e-8, max_iters=1000)
        say("problem", i, reps[i].iterations, r1.iterations)
        assert reps[i].converged == 1 and reps[i].iterations == r1.iterations
        assert same(xs[i], x1.numpy())


def _batch_problems(ctx, sides, with_permuted=False):
    import residuum_hip as rh
    pbs, mats, bs = [], [], []
    for i, s in enumerate(sides):
        pb = rh.Problem(ctx, 3, (s, s, s), (1, 1, 1), 2, list(range(6)), ["0"] * 6)
        pb.set_conductivity(1.0, kappa=np.exp(rh.philox_normal(20261202, i, pb.n_elem)))
        pbs.append(pb)
    if with_permuted:
        pbs.append(permuted_problem(ctx, 9))
    for i, pb in enumerate(pbs):
        K = pb.matrix()
        pb.assemble_matrix(K, [(O.DIFFUSION, 1.0)])
        mats.append(K)
        bs.append(np.random.default_rng(30 + i).standard_normal(pb.n_free))
    return pbs, mats, bs


def test_batched_mixed_bitwise(ctx):
    """The batched mixed-precision solve: each problem's solution, inner
    iteration count, outer cycle count and residuals are bitwise its
    standalone mixed solve (FP32 half storage, uniform stencil kernel) under
    three schedules, and agree with its FP64 solve to the tolerance."""
    import residuum_hip as rh
    pbs, mats, bs = _batch_problems(ctx, [4, 7, 3, 6, 5])
    batch = rh.Batch(ctx, mats)
    db = ctx.from_numpy(batch.pack(bs))
    ref = []
    for pb, K, b in zip(pbs, mats, bs):
        x1 = pb.vector(0.0)
        r1, c1 = K.cg_mixed(ctx.from_numpy(b), x1, tol=1e-8, max_iters=2000)
        x2 = pb.vector(0.0)
        K.cg(ctx.from_numpy(b), x2, tol=1e-12, max_iters=4000)
        ref.append((r1, c1, x1.numpy(), x2.numpy()))
    for sched in ((8, 9, 0, 1), (4, 4, 1, 0), (32, 9, 1, 1)):
        batch.schedule = sched
        dx = ctx.from_numpy(np.zeros(batch.total_rows))
        reps, cyc = batch.cg_mixed(db, dx, tol=1e-8, max_iters=2000)
        xs = batch.unpack(dx.numpy())
        say("schedule", sched, "iterations", [r.iterations for r in reps], "cycles", cyc)
        for i, (r1, c1, x1, x64) in enumerate(ref):
            assert reps[i].converged == r1.converged == 1
            assert reps[i].iterations == r1.iterations and cyc[i] == c1 and c1 >= 2
            assert reps[i].final_residual == r1.final_residual and reps[i].initial_residual == r1.initial_residual
            assert np.array_equal(bits(xs[i]), bits(x1))
            assert np.max(np.abs(xs[i] - x64)) <= 1e-6 * np.max(np.abs(x64))


def test_batched_mixed_explicit_columns(ctx):
    """Renumbered problems (no shared offset patterns, so the FP32 copy
    keeps the sliced ELL layout) batch among themselves bitwise their
    standalone mixed solves; a batch mixing them with half-storage problems
    is refused."""
    import residuum_hip as rh
    pbs = [permuted_problem(ctx, 9), permuted_problem(ctx, 4)]
    mats, bs = [], []
    for i, pb in enumerate(pbs):
        K = pb.matrix()
        pb.assemble_matrix(K, [(O.DIFFUSION, 1.0)])
        mats.append(K)
        bs.append(np.random.default_rng(40 + i).standard_normal(pb.n_free))
    batch = rh.Batch(ctx, mats)
    dx = ctx.from_numpy(np.zeros(batch.total_rows))
    reps, cyc = batch.cg_mixed(ctx.from_numpy(batch.pack(bs)), dx, tol=1e-8, max_iters=2000)
    xs = batch.unpack(dx.numpy())
    for i, (pb, K, b) in enumerate(zip(pbs, mats, bs)):
        x1 = pb.vector(0.0)
        r1, c1 = K.cg_mixed(ctx.from_numpy(b), x1, tol=1e-8, max_iters=2000)
        assert reps[i].iterations == r1.iterations and cyc[i] == c1 and reps[i].converged == 1
        assert np.array_equal(bits(xs[i]), bits(x1.numpy()))
    _, mats2, _ = _batch_problems(ctx, [4])
    mixed = rh.Batch(ctx, mats2 + mats[:1])
    dz = ctx.from_numpy(np.zeros(mixed.total_rows))
    with pytest.raises(rh.RsdError, match="FP32 layouts differ"):
        mixed.cg_mixed(ctx.from_numpy(np.ones(mixed.total_rows)), dz)


@pytest.mark.parametrize("with_permuted", [False, True])
def test_batched_schedules_bitwise(ctx, with_permuted):
    """Every schedule (chunk, load batch, block order, XCD remap) gives the
    same bits, and each problem equals its standalone solve; with a
    renumbered problem in the batch the row products use explicit columns."""
    import residuum_hip as rh
    pbs, mats, bs = _batch_problems(ctx, [4, 7, 3, 6], with_permuted)
    batch = rh.Batch(ctx, mats)
    db = ctx.from_numpy(batch.pack(bs))
    ref = []
    for pb, K, b in zip(pbs, mats, bs):
        x1 = pb.vector(0.0)
        r1 = K.cg(ctx.from_numpy(b), x1, tol=1e-9, max_iters=2000)
        ref.append((r1.iterations, x1.numpy()))
    for chunk in (1, 4, 8, 16, 32):
        for unroll in (4, 9):
            for order in (0, 1):
                for xcd in (0, 1):
                    batch.schedule = (chunk, unroll, order, xcd)
                    assert batch.schedule == (chunk, unroll, order, xcd)
                    dx = ctx.from_numpy(np.zeros(batch.total_rows))
                    reps = batch.cg(db, dx, tol=1e-9, max_iters=2000)
                    xs = batch.unpack(dx.numpy())
                    for i, (its, x1) in enumerate(ref):
                        assert reps[i].converged == 1 and reps[i].iterations == its, (chunk, unroll, order, xcd, i)
                        assert same(xs[i], x1), (chunk, unroll, order, xcd, i)


def test_batched_autotune_cache(ctx, tmp_path):
    """The autotuner times the 32 schedules, keeps the fastest, writes it to
    the JSON cache; a second batch of the same shape takes it from there."""
    import json
    import residuum_hip as rh
    pbs, mats, bs = _batch_problems(ctx, [5, 8, 6])
    cache = str(tmp_path / "tune.json")
    batch = rh.Batch(ctx, mats)
    sched, us, hit = batch.autotune(cache=cache, iters=16)
    say("autotune", sched, us, hit)
    assert not hit and us > 0 and batch.schedule == sched
    d = json.load(open(cache))
    assert len(d) == 1
    (key, entry), = d.items()
    assert "/np=3/" in key and (entry["chunk"], entry["unroll"], entry["order"], entry["xcd"]) == sched
    batch2 = rh.Batch(ctx, mats)
    sched2, us2, hit2 = batch2.autotune(cache=cache)
    assert hit2 and sched2 == sched and batch2.schedule == sched
    # the tuned schedule solves bitwise like the default one
    db = ctx.from_numpy(batch.pack(bs))
    dx1, dx2 = ctx.from_numpy(np.zeros(batch.total_rows)), ctx.from_numpy(np.zeros(batch.total_rows))
    batch.cg(db, dx1, tol=1e-9)
    batch2.schedule = (8, 9, 0, 1)
    batch2.cg(db, dx2, tol=1e-9)
    assert same(dx1.numpy(), dx2.numpy())
    with pytest.raises(rh.RsdError):
        batch.schedule = (8, 5, 0, 1)
    # a cache entry outside the accepted ranges is dropped and tuned again
    entry.update(order=7, chunk=0)
    json.dump({key: entry}, open(cache, "w"))
    batch3 = rh.Batch(ctx, mats)
    sched3, us3, hit3 = batch3.autotune(cache=cache, iters=16)
    assert not hit3 and batch3.schedule == sched3 and sched3[2] in (0, 1) and 1 <= sched3[0] <= 256
    assert json.load(open(cache))[key]["order"] in (0, 1)


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_problem_from_pmsh(ctx, tmp_path):
    """A problem on the reference's mesh file (HeatDispatcher.cpp:48 reads
    the mesh this way) equals the same block problem built on the device,
    and solves the bilinear case exactly; a hex mesh written by the library
    reads back into the same problem."""
    import residuum_hip as rh
    g = "(1-x)*(1-y)"
    pf = rh.Problem.from_pmsh(ctx, os.path.join(GOLDEN, "square_n16.pmsh"), 2, [0, 1, 2, 3], [g] * 4)
    pbk = rh.Problem(ctx, 2, (16, 16), (1, 1), 2, [0, 1, 2, 3], [g] * 4, ext=(-1.0, 1.0, -1.0, 1.0))
    assert (pf.n_nodes, pf.n_elem, pf.n_free, pf.nnz) == (pbk.n_nodes, pbk.n_elem, pbk.n_free, pbk.nnz)
    for a, b in zip(pf.pattern() + pf.dofs(), pbk.pattern() + pbk.dofs()):
        assert same(a, b)
    Kf, Kb = pf.matrix(), pbk.matrix()
    pf.assemble_matrix(Kf, [(O.DIFFUSION, 1.0)])
    pbk.assemble_matrix(Kb, [(O.DIFFUSION, 1.0)])
    assert same(Kf.values(), Kb.values())
    F = pf.vector()
    pf.assemble_load(F)
    pf.apply_essential(F, [(O.DIFFUSION, 1.0)])
    x = pf.vector(0.0)
    rep = Kf.cg(F, x, tol=1e-12, max_iters=5000)
    xyz, _, _ = pf.mesh()
    _, a2t, _ = pf.dofs()
    xy = xyz.reshape(-1, 2)[a2t]
    assert rep.converged and np.max(np.abs(x.numpy() - (1 - xy[:, 0]) * (1 - xy[:, 1]))) <= 1e-10
    # hex mesh through the writer
    xyz3, ien3, rng3 = O.block_mesh(3, (3, 2, 4), (0.0, 1.0, 0.0, 2.0, 0.0, 1.0), (1, 1, 1))
    path = str(tmp_path / "cube.pmsh")
    rh.write_pmsh(path, dict(pdim=3, sdim=3, family=rh.PMSH_HEX, n_nodes=xyz3.size // 3, n_elem=ien3.size // 8,
                             npe=8, faces=6, orders=[1, 1, 1], xyz=xyz3, ien=ien3, rng=rng3))
    ph = rh.Problem.from_pmsh(ctx, path, 2, [4], ["1"])
    ph2 = rh.Problem(ctx, 3, (3, 2, 4), (1, 1, 1), 2, [4], ["1"], ext=(0.0, 1.0, 0.0, 2.0, 0.0, 1.0))
    for a, b in zip(ph.pattern() + ph.dofs(), ph2.pattern() + ph2.dofs()):
        assert same(a, b)


def test_pndf_write_read(ctx, tmp_path):
    """FieldIO::writeBinary / readBinary.  The reference's 24x24 [-1,1]^2 Q1
    mesh carries the mesh tag 17859353126579886777 in the headers of its
    .pndf files (examples/heateq/steady/file_mode); the values are the free
    vector on free nodes and g(x, t) on constrained ones."""
    import struct
    import residuum_hip as rh
    pb = rh.Problem(ctx, 2, (24, 24), (1, 1), 2, [0, 1, 2, 3], ["x+2*y+t"] * 4, ext=(-1.0, 1.0, -1.0, 1.0))
    assert pb.mesh_tag() == 17859353126579886777
    U = np.random.default_rng(1).standard_normal(pb.n_free)
    path = str(tmp_path / "u.pndf")
    pb.write_pndf(path, ctx.from_numpy(U), time=0.5)
    raw = open(path, "rb").read()
    magic, ver, nn, nd, t, tag = struct.unpack("<IIQIdQ", raw[:36])
    assert (magic, ver, nn, nd, t, tag) == (0x504E4446, 1, pb.n_nodes, 1, 0.5, 17859353126579886777)
    vals = np.frombuffer(raw[36:], "<f8")
    xyz, _, _ = pb.mesh()
    t2a, a2t, _ = pb.dofs()
    xy = xyz.reshape(-1, 2)
    assert same(vals[a2t], U)
    c = t2a < 0
    assert same(vals[c], xy[c, 0] + 2 * xy[c, 1] + 0.5)
    nodal, tt = pb.read_pndf(path)
    assert same(nodal, vals) and tt == 0.5
    other = rh.Problem(ctx, 2, (24, 24), (1, 1), 2, [0], ["0"])
    with pytest.raises(rh.RsdError, match="mesh tag"):
        other.read_pndf(path)
