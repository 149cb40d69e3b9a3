"""The CPU oracle pinned against the reference.

1. Golden vectors written by the reference's own headers
   (tests/golden/make_golden.py): every product must match bit for bit.
2. The closed-form values of the reference's fem-tier tests
   (tests/integration/fem/heateq/cpu/HeatEquationMinimal{2D,3D}.cpp).
3. Mathematics: exact bilinear / trilinear solutions, the Kirchhoff
   closed form, 2x2 known systems, eigenmode decay, adjoint vs finite
   differences.
4. When oracle/_ref/libref.so is present: live comparisons on fresh cases.
"""
import glob
import os

import numpy as np
import pytest

import oracle_py as O

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def case_from(g, **kw):
    dim = int(g["dim"])
    return O.Case(dim, tuple(int(v) for v in g["n"]), tuple(int(v) for v in g["p"]), int(g["nq"]),
                  [int(t) for t in g["tags"]], [O.KIND_EXPR[int(k)] for k in g["kinds"]],
                  ext=tuple(float(v) for v in g["ext"]), **kw)


def same(a, b):
    return np.asarray(a).shape == np.asarray(b).shape and np.array_equal(a, b)


@pytest.fixture(params=GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def golden(request):
    return dict(np.load(request.param, allow_pickle=False))


def test_golden_setup(golden):
    c = case_from(golden)
    assert same(c.xyz, golden["xyz"]) and same(c.ien, golden["ien"]) and same(c.rng, golden["rng"])
    assert same(c.t2a, golden["t2a"]) and same(c.ctag, golden["ctag"])
    assert same(c.rp, golden["row_ptr"]) and same(c.ci, golden["col_idx"])
    assert same(c.map, golden["elem_map"])


def test_golden_assembly(golden):
    c = case_from(golden)
    dim = c.dim
    assert same(c.assemble_matrix([(O.DIFFUSION, 1.0)]), golden["K"])
    assert same(c.assemble_matrix([(O.MASS, 1.0)]), golden["M"])
    be = [(O.DIFFUSION, 1.0), (O.MASS | O.SCALED, 250.0)]
    assert same(c.assemble_matrix(be), golden["A_be"])
    src = O.KIND_EXPR[int(golden["src_kind"])]
    F = c.load(src)
    assert same(F, golden["F_load"])
    assert same(c.load(nodal=golden["nodal"]), golden["F_nodal"])
    assert same(c.lift([(O.DIFFUSION, 1.0)], F), golden["F_lift"])
    assert same(c.assemble_vector([(O.MASS, 1.0)], U=golden["U"]), golden["MU"])
    ck = case_from(golden, kappa=golden["kappa"])
    assert same(ck.assemble_matrix([(O.DIFFUSION, 1.0)]), golden["K_kappa"])
    assert same(ck.assemble_matrix(be), golden["A_be_kappa"])
    assert same(ck.assemble_vector([(O.DIFFUSION, 1.0)], U=golden["U"]), golden["KU_kappa"])
    assert same(ck.lift(be, F), golden["F_lift_be_kappa"])
    tensor = (1.0, 0.2, 0.0, 0.2, 0.1, 0.0, 0.0, 0.0, 1.0) if dim == 3 else (1.0, 0.2, 0.2, 0.1)
    ct = case_from(golden, k_expr="1.0 + 0.5*T", dk_expr="0.5", tensor=tensor)
    assert same(ct.assemble_matrix([(O.DIFFUSION, 1.0), (O.TDIFFUSION, 1.0)], U=golden["U"]), golden["J_tdep"])


def test_golden_solvers(golden):
    ck = case_from(golden, kappa=golden["kappa"])
    A, b = golden["A_be_kappa"], golden["b"]
    for pc in (0, 1):
        x, rep = ck.cg(A, b, tol=1e-10, max_iters=2000, precond=pc)
        assert rep.iterations == int(golden[f"cg_pc{pc}_iters"])
        assert same(x, golden[f"cg_pc{pc}_x"])
        x, rep = ck.gmres(A, b, tol=1e-10, max_iters=2000, krylov=7, precond=pc)
        assert rep.iterations == int(golden[f"gmres_pc{pc}_iters"])
        assert same(x, golden[f"gmres_pc{pc}_x"])
    out = ck.be_run(0.0, 0.05, 0.01, cg_tol=1e-9)
    assert out["steps"] == int(golden["be_steps"])
    assert same(out["iters"], golden["be_iters"])
    assert same(out["U"], golden["be_U"])
    if "newton_U" in golden:
        cz = O.Case(ck.dim, ck.n, ck.p, ck.nq, ck.tags, ["0"] * len(ck.tags), ext=ck.ext, k_expr="1.0 + 0.5*T",
                    dk_expr="0.5", tensor=(1.0, 0.2, 0.2, 0.1))
        U, its, _, lin = cz.newton(src=O.KIND_EXPR[3], abs_tol=1e-10, rel_tol=1e-10, lin_solver=1, lin_tol=1e-12,
                                   krylov=30)
        assert its == int(golden["newton_iters"])
        assert same(lin, golden["newton_lin_iters"])
        assert same(U, golden["newton_U"])


# ---------------------------------------------------------------------------
# closed forms from the reference's fem-tier tests
# ---------------------------------------------------------------------------
def dense(c, val):
    A = np.zeros((c.n_free, c.n_free))
    for i in range(c.n_free):
        A[i, c.ci[c.rp[i]:c.rp[i + 1]]] = val[c.rp[i]:c.rp[i + 1]]
    return A


def test_fem2d_closed_forms():
    """HeatEquationMinimal2D.cpp:154-398 (DOF maps, K entries, F, lift)."""
    c = O.Case(2, (4, 4), (1, 1), 2, [0, 1, 3], ["x"] * 3, ext=(0.0, 4.0, 0.0, 4.0))
    assert c.n_nodes == 25 and c.n_free == 12
    expect = {1: 0, 2: 1, 3: 2, 6: 3, 7: 4, 8: 5, 11: 6, 12: 7, 13: 8, 16: 9, 17: 10, 18: 11}
    for node in range(25):
        assert c.t2a[node] == expect.get(node, -1)
    assert c.ctag[10] == 0 and c.ctag[15] == 0 and c.ctag[21] == 3 and c.ctag[9] == 1
    K = dense(c, c.assemble_matrix([(O.DIFFUSION, 1.0)]))
    tol = 1e-10
    assert np.allclose(np.diag(K)[:3], 4 / 3, atol=tol) and np.allclose(np.diag(K)[3:], 8 / 3, atol=tol)
    for i in range(2):
        assert abs(K[i, i + 1] + 1 / 6) < tol
    for i in (3, 6, 7, 9, 10):
        assert abs(K[i, i + 1] + 1 / 3) < tol
    for i in range(9):
        assert abs(K[i, i + 3] + 1 / 3) < tol
    F = c.load("x*y")
    assert np.allclose(F, [1 / 6, 1 / 3, 1 / 2, 1, 2, 3, 2, 4, 6, 3, 6, 9], atol=tol)
    lift = c.lift([(O.DIFFUSION, 1.0)], F) - F
    assert np.allclose(lift, [0, 0, 2, 0, 0, 4, 0, 0, 4, 1, 2, 17 / 3], atol=tol)


def test_fem3d_closed_forms():
    """HeatEquationMinimal3D.cpp:168-236 (constraints and K entries)."""
    c = O.Case(3, (4, 4, 4), (1, 1, 1), 2, [0, 1, 2, 3], ["x"] * 4, ext=(0.0, 4.0) * 3)
    assert c.n_nodes == 125 and c.n_free == 45

    def nid(i, j, k):
        return 25 * k + 5 * j + i
    assert c.t2a[nid(2, 2, 2)] >= 0 and c.t2a[nid(0, 2, 2)] < 0 and c.t2a[nid(4, 2, 2)] < 0
    assert c.ctag[nid(0, 2, 2)] == 0 and c.ctag[nid(4, 2, 2)] == 1
    assert c.ctag[nid(2, 0, 2)] == 2 and c.ctag[nid(2, 4, 2)] == 3
    K = dense(c, c.assemble_matrix([(O.DIFFUSION, 1.0)]))
    a = c.t2a[nid(2, 2, 2)]
    assert abs(K[a, a] - 8 / 3) < 1e-10
    assert abs(K[a, c.t2a[nid(3, 2, 2)]]) < 1e-10
    assert abs(K[a, c.t2a[nid(3, 3, 2)]] + 1 / 6) < 1e-10
    assert abs(K[a, c.t2a[nid(3, 3, 3)]] + 1 / 12) < 1e-10


# ---------------------------------------------------------------------------
# mathematics
# ---------------------------------------------------------------------------
def test_config1_exact_bilinear():
    """64x64 Q1 with g=(1-x)(1-y): the discrete solution is the nodal
    interpolant (HeatEquationMinimal2D full-tier test, config 1)."""
    c = O.Case(2, (64, 64), (1, 1), 2, [0, 1, 2, 3], ["(1-x)*(1-y)"] * 4)
    U, rep = c.steady(tol=1e-10, max_iters=5000)
    assert rep.converged
    xy = c.xyz.reshape(-1, 2)[c.a2t]
    exact = (1 - xy[:, 0]) * (1 - xy[:, 1])
    assert np.max(np.abs(U - exact)) < 1e-8


def test_trilinear_exact_hex():
    c = O.Case(3, (8, 8, 8), (1, 1, 1), 2, list(range(6)), ["x+2*y+3*z"] * 6)
    U, rep = c.steady(tol=1e-12, max_iters=5000)
    xyz = c.xyz.reshape(-1, 3)[c.a2t]
    assert np.max(np.abs(U - (xyz[:, 0] + 2 * xyz[:, 1] + 3 * xyz[:, 2]))) < 1e-9


def test_kirchhoff_newton():
    """HeatEquationSteadyNonlinearConductivity2D.cpp:29-64: k(T)=1+0.5T,
    T=0 at x=0, T=1 at x=1; theta(x) = (-1 + sqrt(1 + 2*0.5*u(x)))/0.5."""
    c = O.Case(2, (40, 4), (1, 1), 2, [0, 1], ["0.0", "1.0"], k_expr="1.0 + 0.5*T", dk_expr="0.5")
    U, its, res, _ = c.newton(abs_tol=1e-10, rel_tol=1e-10, lin_solver=0, lin_tol=1e-12, lin_max=5000)
    assert its > 0
    x = c.xyz.reshape(-1, 2)[c.a2t][:, 0]
    u = (1.0 + 0.25) * x  # kirchhoff(1) * s
    theta = (-1.0 + np.sqrt(1.0 + 2 * 0.5 * u)) / 0.5
    assert np.max(np.abs(U - theta)) < 1e-6


def test_q2_newton_gmres30_growth():
    """Restarted GMRES(30) + Jacobi on the Newton Jacobian of the Q2
    anisotropic k(T) problem (k = (1 + 0.5 T) diag(1, 0.1), f = 10 sin(pi x)
    sin(pi y), T = 0 on the boundary) after one Newton step, with that
    step's residual as right-hand side.  The Arnoldi steps to a fixed
    relative tolerance grow 2.7-4.4x per mesh halving (the condition
    number grows 4x and restarting discards the Krylov space), which is what
    makes an inexact inner tolerance necessary at 512^2 (~1.05M DOFs), where
    1e-3 already takes ~30k steps.  The 64^2 counts are the ones the device
    GMRES reports in the GPU reduction order."""
    src = "10*sin(pi*x)*sin(pi*y)"
    steps = {}
    for n in (16, 32, 64):
        c = O.Case(2, (n, n), (2, 2), 3, [0, 1, 2, 3], ["0"] * 4, k_expr="1.0 + 0.5*T", dk_expr="0.5",
                   tensor=(1.0, 0.0, 0.0, 0.1), order=O.ORDER_GPU)
        U, _, _, _ = c.newton(src=src, abs_tol=0.0, rel_tol=0.0, max_newton=1, lin_solver=1, lin_tol=1e-10,
                              lin_max=100000)
        J = c.assemble_matrix([(O.DIFFUSION, 1.0), (O.TDIFFUSION, 1.0)], U=U)
        R = c.load(src=src) - c.assemble_vector([(O.DIFFUSION, 1.0)], U=U)
        for tol in (1e-4, 1e-8):
            x, rep = c.gmres(J, R, tol=tol, max_iters=100000, krylov=30)
            assert rep.converged and rep.final_residual_rel < tol
            steps[n, tol] = rep.iterations
    print(steps)
    assert steps[64, 1e-4] == 1168
    for tol in (1e-4, 1e-8):
        assert steps[32, tol] > 2.4 * steps[16, tol] and steps[64, tol] > 4.0 * steps[32, tol]


def test_cg_gmres_2x2_known():
    """tests/unit/linalg/solver/iterative/cg/Solver.cpp: a 2x2 SPD system."""
    c = O.Case(2, (1, 1), (1, 1), 2, [], [])  # 4 free nodes, a full 4x4 pattern
    A = np.array([[4.0, 1, 0, 0], [1, 3, 0, 0], [0, 0, 2, 0], [0, 0, 0, 5]])
    val = np.array([A[i, j] for i in range(4) for j in c.ci[c.rp[i]:c.rp[i + 1]]])
    b = np.array([1.0, 2.0, 3.0, 4.0])
    for order in (O.ORDER_REF, O.ORDER_GPU):
        x, rep = c.cg(val, b, tol=1e-14, order=order)
        assert np.allclose(x, np.linalg.solve(A, b), atol=1e-13)
        x, rep = c.gmres(val, b, tol=1e-14, krylov=2, order=order)
        assert np.allclose(x, np.linalg.solve(A, b), atol=1e-13)


def test_jacobi_inverse_diag():
    c = O.Case(2, (3, 3), (1, 1), 2, [0], ["0"])
    val = c.assemble_matrix([(O.DIFFUSION, 1.0)])
    d = np.zeros(c.n_free)
    O.lib().orc_inv_diag(O.C.c_int64(c.n_free), O.lp(c.rp), O.lp(c.ci), O.dp(val), O.dp(d))
    assert np.allclose(d, 1.0 / np.diag(dense(c, val)))


def test_gpu_order_reduction_is_a_sum():
    rng = np.random.default_rng(0)
    # block edges and a size with more than 1024 blocks (fold lane wrap-around)
    for n in (0, 1, 255, 256, 257, 70000, 300001):
        x, y = rng.standard_normal(n), rng.standard_normal(n)
        assert abs(O.dot(x, y, O.ORDER_GPU) - np.dot(x, y)) <= 1e-12 * (np.abs(x * y).sum() + 1e-300)
        assert O.dot(x, y, O.ORDER_REF) == sum((x * y).tolist(), 0.0) or n == 0


def test_be_eigenmode_decay():
    """HeatEquationTransientBackwardEuler2D.cpp:38-46,194-251: the sin-sin
    mode decays like exp(-2 pi^2 t) (within 5%), monotonically."""
    c = O.Case(2, (32, 32), (1, 1), 2, [0, 1, 2, 3], ["0"] * 4)
    xy = c.xyz.reshape(-1, 2)[c.a2t]
    U0 = np.sin(np.pi * xy[:, 0]) * np.sin(np.pi * xy[:, 1])
    out = c.be_run(0.0, 0.05, 1e-3, U0=U0, cg_tol=1e-12, hist=True)
    amp = out["U_hist"].max(axis=1)
    assert np.all(np.diff(amp) < 0)
    assert abs(amp[-1] / np.exp(-2 * np.pi ** 2 * 0.05) - 1.0) < 0.05


def test_be_linear_field_preserved():
    """BackwardEulerStage.cpp:278-354: a linear steady field stays exact."""
    c = O.Case(3, (4, 4, 4), (1, 1, 1), 2, list(range(6)), ["x+2*y+3*z"] * 6)
    xyz = c.xyz.reshape(-1, 3)[c.a2t]
    U0 = xyz[:, 0] + 2 * xyz[:, 1] + 3 * xyz[:, 2]
    out = c.be_run(0.0, 0.03, 0.01, U0=U0, cg_tol=1e-12)
    assert np.max(np.abs(out["U"] - U0)) < 1e-8


def test_adjoint_matches_finite_differences():
    """The discrete adjoint (no reference counterpart) against central
    differences on a small hex problem."""
    c = O.Case(3, (4, 4, 4), (1, 1, 1), 2, [4], ["1"])
    theta = 0.5 * O.normal_field(7, 0, c.n_elem)
    obs = np.arange(0, c.n_free, 7)
    tobs = np.linspace(0.1, 0.5, obs.size)
    c.kappa = np.exp(theta)
    c.set_order(O.ORDER_REF)
    loss, grad, _ = c.adjoint(0.0, 5e-3, 4, obs, tobs)
    rng = np.random.default_rng(1)
    h = 1e-6
    for _ in range(3):
        d = rng.standard_normal(c.n_elem)
        ls = []
        for s in (1, -1):
            c.kappa = np.exp(theta + s * h * d)
            c.set_order(O.ORDER_REF)
            ls.append(c.adjoint(0.0, 5e-3, 4, obs, tobs)[0])
        fd = (ls[0] - ls[1]) / (2 * h)
        assert abs(fd - grad @ d) <= 1e-6 * max(abs(fd), 1e-12)


def test_adjoint_scaled_conductivity_and_repeated_observation():
    """The adjoint with a conductivity value k = 2 (the element matrices carry
    kappa_e * k) and an observation index listed twice (the loss counts it
    twice): central differences of the loss itself pin both."""
    c = O.Case(3, (3, 4, 3), (1, 1, 1), 2, [4], ["1"])
    c.k_scalar = 2.0
    theta = 0.4 * O.normal_field(11, 0, c.n_elem)
    obs = np.concatenate([np.arange(1, c.n_free, 5), [3, 3]])
    tobs = np.linspace(0.05, 0.4, obs.size)
    c.kappa = np.exp(theta)
    c.set_order(O.ORDER_REF)
    loss, grad, UN = c.adjoint(0.0, 5e-3, 3, obs, tobs)
    assert np.isclose(loss, 0.5 * np.sum((UN[obs] - tobs) ** 2), rtol=1e-14)
    rng = np.random.default_rng(3)
    h = 1e-6
    for _ in range(3):
        d = rng.standard_normal(c.n_elem)
        ls = []
        for s_ in (1, -1):
            c.kappa = np.exp(theta + s_ * h * d)
            c.set_order(O.ORDER_REF)
            ls.append(c.adjoint(0.0, 5e-3, 3, obs, tobs)[0])
        fd = (ls[0] - ls[1]) / (2 * h)
        assert abs(fd - grad @ d) <= 1e-6 * max(abs(fd), 1e-12)


# ---------------------------------------------------------------------------
# live comparison with the reference harness (only where it was built)
# ---------------------------------------------------------------------------
@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref/libref.so not built (needs /root/reference)")
@pytest.mark.parametrize("spec", [
    (2, (6, 4), (1, 1), 2, [0, 2], [7, 7], None),
    (2, (3, 4), (2, 2), 3, [1, 3], [2, 2], None),
    (3, (3, 2, 4), (1, 1, 1), 2, [4, 5], [6, 0], None),
    (3, (2, 3, 2), (2, 2, 2), 3, [0], [7], None),
])
def test_live_reference(spec):
    dim, n, p, nq, tags, kinds, ext = spec
    c = O.Case(dim, n, p, nq, tags, [O.KIND_EXPR[k] for k in kinds], ext=ext)
    r = O.RefCase(dim, n, p, nq, tags, kinds, ext=ext)
    xyz, ien, rng = r.mesh()
    assert same(c.xyz, xyz) and same(c.ien, ien) and same(c.rng, rng)
    t2a, ct = r.dofs()
    assert same(c.t2a, t2a) and same(c.ctag, ct)
    rp, ci = r.pattern()
    assert same(c.rp, rp) and same(c.ci, ci) and same(c.map, r.map())
    kap = np.exp(np.random.default_rng(2).standard_normal(c.n_elem))
    c.kappa = kap
    c.set_order(O.ORDER_REF)
    A = c.assemble_matrix([(O.DIFFUSION, 1.0), (O.MASS | O.SCALED, 40.0)])
    assert same(A, r.assemble_matrix(2, 40.0, kappa=kap))
    out = c.be_run(0.0, 0.03, 0.01, cg_tol=1e-9)
    Ur, st, its = r.be_run(0.0, 0.03, 0.01, tol=1e-9, kappa=kap)
    assert out["steps"] == st and same(out["iters"], its) and same(out["U"], Ur)


# ---------------------------------------------------------------------------
# natural (flux) boundary conditions (BoundaryApplicator.tpp:66-140)
# ---------------------------------------------------------------------------
NATURAL_CASES = [
    # mesh, Dirichlet tags / kinds, natural entries (tag, flux kind or "nodal")
    ((2, (6, 5), (1, 1), 2, [0], [7], (0.0, 2.0, -1.0, 1.0)), [(1, 0), (2, 1), (3, 2), (2, "nodal")]),
    ((2, (4, 3), (2, 2), 3, [3], [6], None), [(0, 1), (1, 2), (2, "nodal"), (0, 0)]),
    ((3, (3, 4, 2), (1, 1, 1), 2, [4], [6], None), [(0, 2), (1, 0), (5, "nodal"), (2, 1), (3, 2)]),
    ((3, (2, 2, 3), (1, 1, 1), 2, [], [], None), [(4, 1), (5, 2), (0, "nodal")]),
]


def _natural_entries(c, spec, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for tag, fl in spec:
        out.append((tag, rng.standard_normal(c.n_nodes * c.dim) if fl == "nodal" else fl))
    return out


@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref/libref.so not built (needs /root/reference)")
@pytest.mark.parametrize("case", NATURAL_CASES)
def test_natural_bc_live_reference(case):
    """The oracle's natural BC terms bit for bit against the reference's own
    applyNaturalBCs with FluxBoundaryForm / NodalFluxForm entries (several
    entries on one tag, a tag with no flux, a constrained face among them)."""
    (dim, n, p, nq, tags, kinds, ext), spec = case
    c = O.Case(dim, n, p, nq, tags, [O.KIND_EXPR[k] for k in kinds], ext=ext)
    r = O.RefCase(dim, n, p, nq, tags, kinds, ext=ext)
    entries = _natural_entries(c, spec)
    F0 = np.random.default_rng(9).standard_normal(c.n_free)
    for time in (0.0, 0.7):
        got = c.apply_natural(entries, F0, time=time)
        want = r.natural(entries, F0, time=time)
        assert same(got, want)
        assert not np.array_equal(got, F0)


def test_natural_bc_closed_forms():
    """Flux integrals the quadrature computes exactly: the free and the
    constrained nodes' face terms add up to the boundary integral of q.n.
    Unit square, no Dirichlet faces, so every node is free: constant
    q = (1.5, -0.5) on x = 0 gives -1.5; q = (x, y^2) on x = 1 gives 1 and on
    y = 1 gives 1; on the unit cube q = (x, y^2, x z) on z = 1 (tag 5) gives
    the integral of x over the face, 1/2."""
    c = O.Case(2, (5, 3), (1, 1), 2, [], [])
    z = np.zeros(c.n_free)
    assert np.isclose(c.apply_natural([(0, 0)], z).sum(), -1.5, rtol=0, atol=1e-14)
    assert np.isclose(c.apply_natural([(1, 1)], z).sum(), 1.0, rtol=0, atol=1e-14)
    assert np.isclose(c.apply_natural([(3, 1)], z).sum(), 1.0, rtol=0, atol=1e-14)
    # the same flux on two entries of one tag doubles the terms (the two
    # entries interleave per point, so to rounding)
    one = c.apply_natural([(1, 1)], z)
    assert np.allclose(c.apply_natural([(1, 1), (1, 1)], z), 2 * one, rtol=1e-14, atol=0)
    c3 = O.Case(3, (2, 3, 2), (1, 1, 1), 2, [], [])
    z3 = np.zeros(c3.n_free)
    assert np.isclose(c3.apply_natural([(5, 1)], z3).sum(), 0.5, rtol=0, atol=1e-14)
    # the nodal form interpolates: nodal values of a linear flux reproduce it
    xyz = c3.xyz.reshape(-1, 3)
    nod = np.stack([xyz[:, 0], 2.0 + 0 * xyz[:, 0], xyz[:, 0] * 0.5], axis=1).ravel()
    assert np.isclose(c3.apply_natural([(5, nod)], z3).sum(), 0.25, rtol=0, atol=1e-14)
    # Q2 quads: the flux (x, y^2) on y = 1 (tag 3) integrates to 1 as well
    c2 = O.Case(2, (3, 2), (2, 2), 3, [], [])
    assert np.isclose(c2.apply_natural([(3, 1)], np.zeros(c2.n_free)).sum(), 1.0, rtol=0, atol=1e-14)


# ---------------------------------------------------------------------------
# the matrix-free operator (FEMOperator.hpp:27-71) and assembleDiagonal
# ---------------------------------------------------------------------------
DIAG_CASES = [
    (2, (5, 4), (1, 1), 2, [0, 2], ["x", "x"]),
    (2, (3, 3), (2, 2), 3, [1], ["1"]),
    (3, (3, 2, 3), (1, 1, 1), 2, [4], ["1"]),
]
BE_FORMS = [(O.DIFFUSION, 1.0), (O.MASS | O.SCALED, 40.0)]


def diag_slots(c):
    return np.array([c.rp[i] + int(np.searchsorted(c.ci[c.rp[i]:c.rp[i + 1]], i)) for i in range(c.n_free)])


@pytest.mark.parametrize("order", [O.ORDER_REF, O.ORDER_GPU])
@pytest.mark.parametrize("spec", DIAG_CASES)
def test_assemble_diagonal_is_matrix_diagonal(spec, order):
    """assembleDiagonal adds each element's Ke[a][a] in element order, which
    is how assembleMatrix builds the diagonal entries: the two agree bit for
    bit, in either element order."""
    dim, n, p, nq, tags, exprs = spec
    c = O.Case(dim, n, p, nq, tags, exprs)
    c.kappa = np.exp(0.5 * np.random.default_rng(4).standard_normal(c.n_elem))
    c.set_order(order)
    ds = diag_slots(c)
    for forms in ([(O.DIFFUSION, 1.0)], [(O.MASS, 1.0)], BE_FORMS):
        A = c.assemble_matrix(forms)
        assert same(c.assemble_diagonal(forms), A[ds])


@pytest.mark.parametrize("spec", DIAG_CASES)
def test_fem_apply_matches_csr(spec):
    """FEMOperator::apply (assembleVector on the operand) is A x summed
    element by element instead of row by row: equal to rounding."""
    dim, n, p, nq, tags, exprs = spec
    c = O.Case(dim, n, p, nq, tags, exprs)
    c.kappa = np.exp(0.5 * np.random.default_rng(5).standard_normal(c.n_elem))
    c.set_order(O.ORDER_REF)
    A = c.assemble_matrix(BE_FORMS)
    x = np.random.default_rng(6).standard_normal(c.n_free)
    y_fem = c.assemble_vector(BE_FORMS, operand=x, selector=O.SEL_OPERAND)
    y_csr = c.spmv(A, x)
    scale = c.spmv(np.abs(A), np.abs(x))
    assert np.max(np.abs(y_fem - y_csr) / scale) < 1e-14


def test_matrix_free_cg_exact_bilinear():
    """The reference's matrix-free CG test (tests/integration/full/heateq/cpu/
    HeatEquationMinimal2D.cpp:249-320): 24x24 Q1 on [-1,1]^2, g = (1-x)(1-y)
    on all sides, no source, CG + Identity to 1e-12: the nodal solution is
    (1-x)(1-y) to 1e-10 and the final residual is 1e-12 +- 1e-11."""
    g = "(1-x)*(1-y)"
    c = O.Case(2, (24, 24), (1, 1), 2, [0, 1, 2, 3], [g] * 4, ext=(-1.0, 1.0, -1.0, 1.0))
    assert c.n_free == 529
    F = c.lift([(O.DIFFUSION, 1.0)], c.load())
    x, rep = c.cg_fem([(O.DIFFUSION, 1.0)], F, tol=1e-12, max_iters=10000, precond=0)
    assert rep.converged == 1 and abs(rep.final_residual - 1e-12) <= 1e-11
    xy = c.xyz.reshape(-1, 2)[c.a2t]
    assert np.max(np.abs(x - (1 - xy[:, 0]) * (1 - xy[:, 1]))) <= 1e-10
    xc, repc = c.cg(c.assemble_matrix([(O.DIFFUSION, 1.0)]), F, tol=1e-12, max_iters=10000, precond=0)
    assert repc.converged == 1 and np.max(np.abs(x - xc)) <= 1e-10


@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref/libref.so not built (needs /root/reference)")
@pytest.mark.parametrize("spec", [
    (2, (6, 4), (1, 1), 2, [0, 2], [7, 7], None),
    (2, (3, 4), (2, 2), 3, [1, 3], [2, 2], None),
    (3, (3, 2, 4), (1, 1, 1), 2, [4, 5], [6, 0], None),
])
def test_live_reference_matrix_free(spec):
    """assembleDiagonal and cg::Solver over FEMOperator, oracle against the
    reference's own templates: bit for bit, iteration counts equal."""
    dim, n, p, nq, tags, kinds, ext = spec
    c = O.Case(dim, n, p, nq, tags, [O.KIND_EXPR[k] for k in kinds], ext=ext)
    r = O.RefCase(dim, n, p, nq, tags, kinds, ext=ext)
    kap = np.exp(np.random.default_rng(2).standard_normal(c.n_elem))
    c.kappa = kap
    c.set_order(O.ORDER_REF)
    assert same(c.assemble_diagonal([(O.DIFFUSION, 1.0)]), r.assemble_diagonal(0, kappa=kap))
    assert same(c.assemble_diagonal([(O.MASS, 1.0)]), r.assemble_diagonal(1, kappa=kap))
    assert same(c.assemble_diagonal(BE_FORMS), r.assemble_diagonal(2, 40.0, kappa=kap))
    b = np.random.default_rng(3).standard_normal(c.n_free)
    for pc in (0, 1):
        x, rep = c.cg_fem(BE_FORMS, b, tol=1e-10, max_iters=500, precond=pc)
        xr, ok, it = r.cg_fem(2, b, coef=40.0, kappa=kap, precond=pc, tol=1e-10, max_iters=500)
        assert ok and rep.converged == 1 and rep.iterations == it and same(x, xr)


# ---------------------------------------------------------------------------
# weak boundary reaction (QuantityEvaluator.tpp:95-199)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dim,n,p,ext,a_axis", [
    (2, (4, 4), (1, 1), (-1.0, 1.0, -1.0, 1.0), 0),
    (2, (3, 5), (2, 2), (0.0, 2.0, 0.0, 1.0), 1),
    (3, (3, 2, 2), (1, 1, 1), (0.0, 1.0, 0.0, 2.0, 0.0, 3.0), 2),
])
def test_boundary_reaction_linear_field(dim, n, p, ext, a_axis):
    """The reference's heat-flux fixture (tests/unit/equation/heateq/
    QuantityEvaluator.cpp:20-100: k = 2, T = a x + b y (+ c z), no
    essential BCs): with no source the weak reaction on the face x_i = min
    is k * dT/dx_i times the face's measure (the corner terms of the
    adjacent faces cancel), and the measure is the face's length/area."""
    c = O.Case(dim, n, p, p[0] + 1, [], [], ext=ext, k_scalar=2.0)
    xyz = c.xyz.reshape(-1, dim)
    grad = [1.0, 2.0, -0.5][:dim]
    U = (xyz[c.a2t] @ np.array(grad))
    tag = 2 * a_axis  # the face at the low end of that axis
    res, meas = c.boundary_reaction([tag], U)
    lens = [ext[2 * i + 1] - ext[2 * i] for i in range(dim)]
    area = float(np.prod([lens[i] for i in range(dim) if i != a_axis]))
    assert abs(meas[0] - area) < 1e-12 * area
    assert abs(res[0] - 2.0 * grad[a_axis] * area) < 1e-12 * max(1.0, abs(2.0 * grad[a_axis] * area))


@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref/libref.so not built (needs /root/reference)")
@pytest.mark.parametrize("spec", [
    (2, (6, 4), (1, 1), 2, [0, 2], [7, 7], None),
    (2, (3, 4), (2, 2), 3, [1, 3], [2, 2], None),
    (3, (3, 2, 4), (1, 1, 1), 2, [4, 5], [6, 0], None),
])
def test_live_reference_boundary_reaction(spec):
    """evaluateBoundaryReaction with the weak heat-flux pairs, oracle against
    the reference's own QuantityEvaluator: bit for bit, steady and
    transient, Dirichlet and free tags together."""
    dim, n, p, nq, tags, kinds, ext = spec
    c = O.Case(dim, n, p, nq, tags, [O.KIND_EXPR[k] for k in kinds], ext=ext)
    r = O.RefCase(dim, n, p, nq, tags, kinds, ext=ext)
    rng = np.random.default_rng(11)
    U, Ud = rng.standard_normal(c.n_free), rng.standard_normal(c.n_free)
    all_tags = list(range(2 * dim))
    for src_kind in (0, 5 if dim == 3 else 4):
        for Udot in (None, Ud):
            got = c.boundary_reaction(all_tags, U, Udot=Udot, src=O.KIND_EXPR[src_kind], time=0.25)
            want = r.boundary_reaction(all_tags, U, Udot=Udot, src_kind=src_kind, time=0.25)
            assert same(got[0], want[0]) and same(got[1], want[1])


# ---------------------------------------------------------------------------
# strong boundary heat flux (QuantityEvaluator.tpp:14-93, HeatFluxIntegrand)
# ---------------------------------------------------------------------------
def _flux_fixture(dirichlet=False):
    """tests/unit/equation/heateq/QuantityEvaluator.cpp:33-95: 4x4 Q1 on
    [-1,1]^2, k = 2, T = x + 2y, two-point rules; no essential BCs, or (the
    constrained case, :139-183) T as a Dirichlet value on all four sides."""
    tags = [0, 1, 2, 3] if dirichlet else []
    c = O.Case(2, (4, 4), (1, 1), 2, tags, ["x + 2*y"] * len(tags), ext=(-1.0, 1.0, -1.0, 1.0), k_scalar=2.0)
    xy = c.xyz.reshape(-1, 2)[c.a2t]
    return c, xy[:, 0] + 2.0 * xy[:, 1]


def test_flux_average_left_boundary():
    """HeatFluxAverageOnLeftBoundaryMatchesAnalytic (:101-121): k a."""
    c, U = _flux_fixture()
    res, meas = c.boundary_flux([0], U)
    assert abs(res[0] / meas[0] - 2.0) <= 1e-10


def test_flux_integral_and_length():
    """MultipleQuantitiesInOnePassAreIndependentlyCorrect (:123-145)."""
    c, U = _flux_fixture()
    res, meas = c.boundary_flux([0], U)
    assert abs(res[0] - 4.0) <= 1e-10 and abs(meas[0] - 2.0) <= 1e-10


def test_flux_target_dirichlet_constrained():
    """HeatFluxIntegralIsCorrectWhenTargetBoundaryIsAlsoDirichletConstrained
    (:147-190): the corner values come from the constrained gather."""
    c, U = _flux_fixture(dirichlet=True)
    res, _ = c.boundary_flux([0], U)
    assert abs(res[0] - 4.0) <= 1e-10


def test_flux_registry_and_combinations():
    """BoundaryQuantityRegistryAndCombinationMatchAnalytic (:192-238): all
    four sides in one pass, combinations sum coefficient x result, and the
    harmonic field's net flux is zero."""
    c, U = _flux_fixture()
    res, _ = c.boundary_flux([0, 1, 2, 3], U)
    want = np.array([4.0, -4.0, 8.0, -8.0])
    assert np.max(np.abs(res - want)) <= 1e-10
    comb = lambda terms: sum(coef * res[t] for t, coef in terms)
    assert abs(comb([(1, 1.0), (2, 1.0)]) - 4.0) <= 1e-10
    assert abs(comb([(1, 1.0), (3, -1.0)]) - 4.0) <= 1e-10
    assert abs(comb([(t, 1.0) for t in range(4)])) <= 1e-9


def test_weak_reaction_on_the_flux_fixture():
    """The weak reaction of the same harmonic field, no source: equal to the
    strong flux per side (the linear field is reproduced exactly), and the
    four sides sum to zero."""
    c, U = _flux_fixture()
    res, meas = c.boundary_reaction([0, 1, 2, 3], U)
    assert np.max(np.abs(res - np.array([4.0, -4.0, 8.0, -8.0]))) <= 1e-10
    assert np.max(np.abs(meas - 2.0)) <= 1e-14
    assert abs(res.sum()) <= 1e-9


@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref/libref.so not built (needs /root/reference)")
@pytest.mark.parametrize("spec", [
    (2, (6, 4), (1, 1), 2, [0, 2], [7, 7], {}),
    (2, (3, 4), (2, 2), 3, [1, 3], [2, 2], {"tensor": (2.0, 0.5, 0.5, 1.0)}),
    (2, (5, 3), (1, 1), 3, [0], [1], {"tdep": True}),
    (3, (3, 2, 4), (1, 1, 1), 2, [4, 5], [6, 0], {}),
])
def test_live_reference_boundary_flux(spec):
    """evaluateBoundaryRegistry with HeatFluxIntegrand: oracle against the
    reference's own code bit for bit (isotropic, anisotropic tensor,
    k(T) = 1 + 0.5 T; free and Dirichlet tags)."""
    dim, n, p, nq, tags, kinds, model = spec
    kw = {}
    if model.get("tensor"):
        kw["tensor"] = model["tensor"]
    if model.get("tdep"):
        kw.update(k_expr="1 + 0.5*T", dk_expr="0.5")
    c = O.Case(dim, n, p, nq, tags, [O.KIND_EXPR[k] for k in kinds], **kw)
    r = O.RefCase(dim, n, p, nq, tags, kinds)
    if model.get("tensor"):
        r.set_model(tensor=list(model["tensor"]) + [0.0] * (9 - len(model["tensor"])))
    if model.get("tdep"):
        r.set_model(tdep=1, ka=1.0, kb=0.5)
    U = np.random.default_rng(21).standard_normal(c.n_free)
    all_tags = list(range(2 * dim))
    got = c.boundary_flux(all_tags, U, time=0.75)
    want = r.boundary_flux(all_tags, U, time=0.75)
    assert same(got[0], want[0]) and same(got[1], want[1])
