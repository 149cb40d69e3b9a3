"""Full-size runs of the workloads the hot path serves, checked through
properties that do not need a CPU solve of the same size (a CPU solve at
these sizes takes minutes to hours): gradient checks, symmetry, residual
checks, batch-vs-standalone equality.  Timings are printed."""
import time

import numpy as np
import pytest

import oracle_py as O  # noqa: F401  (keeps the import path of the suite uniform)


def say(*a):
    print(*a, flush=True)


@pytest.mark.gpu
def test_adjoint_gradient_96cube(ctx):
    """96^3 Q1 hex, 20 Backward Euler steps of dt = 5e-3, T = 1 on z = 0,
    kappa_e = exp(theta_e) with theta a Gaussian random field (std 0.5,
    correlation length 0.1, seed 20261203); loss 1/2 ||T_N[obs] - T_obs||^2
    over 4096 free nodes (seed 20261204), T_obs from the forward run with
    theta + 0.1 xi (xi: seed 20261205).  The adjoint gradient with respect
    to all 884,736 theta_e is checked against central differences
    (h = 1e-6) along 32 Philox directions (seed 20261206) and by the
    second-order Taylor remainder.

    Tolerance: each direction's |FD - g.d| within 1e-6 of the RMS of g.d
    over the 32 directions (a direction nearly orthogonal to g has g.d close
    to 0, where a per-direction ratio only measures the FD noise floor of
    about 1e-9), and the median per-direction relative difference below
    1e-6."""
    import residuum_hip as rh
    from residuum_hip.fields import gaussian_random_field, sample_without_replacement
    n = 96
    dt, steps, tol = 5e-3, 20, 1e-13
    t0 = time.perf_counter()
    pb = rh.Problem(ctx, 3, (n, n, n), (1, 1, 1), 2, [4], ["1"])
    pb.set_mass(1.0, 1.0)
    say(f"setup {time.perf_counter() - t0:.2f}s n_free={pb.n_free} n_elem={pb.n_elem}")
    theta = gaussian_random_field(20261203, (n, n, n), 0.1, 0.5)
    xi = gaussian_random_field(20261205, (n, n, n), 0.1, 0.5)
    obs = sample_without_replacement(20261204, pb.n_free, 4096)
    U0 = pb.vector(0.0)
    dUN = ctx.empty(pb.n_free)
    dk = ctx.from_numpy(np.exp(theta + 0.1 * xi))
    pb.adjoint(dk, U0, 0.0, dt, steps, np.zeros(0, np.int64), np.zeros(0), tol=tol, want_grad=False, d_UN=dUN)
    t_obs = dUN.numpy()[obs]

    def loss(th):
        dk.upload(np.exp(th))
        return pb.adjoint(dk, U0, 0.0, dt, steps, obs, t_obs, tol=tol, want_grad=False)

    dg = ctx.empty(pb.n_elem)
    dk.upload(np.exp(theta))
    ctx.sync()
    ta = time.perf_counter()
    L = pb.adjoint(dk, U0, 0.0, dt, steps, obs, t_obs, tol=tol, want_grad=True, d_grad=dg)
    ta = time.perf_counter() - ta
    tf = time.perf_counter()
    L2 = loss(theta)
    tf = time.perf_counter() - tf
    g = dg.numpy()
    say(f"loss {L:.6e}; forward {tf*1e3:.0f} ms, forward + adjoint sweep {ta*1e3:.0f} ms")
    assert L2 == L and np.all(np.isfinite(g))
    h = 1e-6
    ads, fds = [], []
    for j in range(32):
        d = rh.philox_normal(20261206, j, pb.n_elem)
        fds.append((loss(theta + h * d) - loss(theta - h * d)) / (2 * h))
        ads.append(float(g @ d))
        say(f"direction {j:2d}: adjoint {ads[-1]:+.10e}  central FD {fds[-1]:+.10e}  "
            f"rel diff {abs(fds[-1] - ads[-1]) / abs(ads[-1]):.2e}")
    ads, fds = np.array(ads), np.array(fds)
    scale = np.sqrt(np.mean(ads ** 2))
    nerr = np.abs(fds - ads) / scale
    rel = np.abs(fds - ads) / np.abs(ads)
    say(f"|FD - g.d| / rms(g.d): max {nerr.max():.2e}; per-direction relative: median {np.median(rel):.2e}")
    # Taylor remainder |L(theta + e d) - L - e g.d| = O(e^2)
    d = rh.philox_normal(20261206, 0, pb.n_elem)
    gd = float(g @ d)
    eps = [4e-2, 2e-2, 1e-2]
    rem = [abs(loss(theta + e * d) - L - e * gd) for e in eps]
    slopes = [np.log2(rem[i] / rem[i + 1]) for i in range(2)]
    say(f"Taylor remainders {rem}, slopes {slopes}")
    assert min(slopes) > 1.8
    assert nerr.max() < 1e-6 and np.median(rel) < 1e-6


@pytest.mark.gpu
def test_batched_256_problems(ctx):
    """256 independent Q1 hex problems on unit cubes of side s ~ U{16..64}
    elements (Philox seed 20261202), kappa_e lognormal(0, 1), nodal source
    U[0, 1], T = 0 on all faces; CG + Jacobi to 1e-8 from x = 0.  One batched
    solve (one set of launches for all problems) against 256 standalone
    FP64 solves: every problem bitwise equal, iteration counts equal.  Then
    the batched mixed-precision solve: every problem within 1e-6 of its FP64
    solution, and every 16th bitwise its standalone mixed solve."""
    import residuum_hip as rh
    u = rh.philox_uniform(20261202, 0, 256)
    sides = (16 + np.floor(u * 49)).astype(int)
    t0 = time.perf_counter()
    pbs, mats, bs = [], [], []
    for i, s in enumerate(sides):
        pb = rh.Problem(ctx, 3, (s, s, s), (1, 1, 1), 2, list(range(6)), ["0"] * 6)
        pb.set_conductivity(1.0, kappa=np.exp(rh.philox_normal(20261202, 1 + i, pb.n_elem)))
        K = pb.matrix()
        pb.assemble_matrix(K, [(rh.DIFFUSION, 1.0)])
        F = pb.vector()
        nod = ctx.from_numpy(rh.philox_uniform(20261202, 1000 + i, pb.n_nodes))
        pb.assemble_load(F, nodal=nod)
        nod.free()
        pbs.append(pb); mats.append(K); bs.append(F.numpy())
        F.free()
    n_tot = sum(pb.n_free for pb in pbs)
    nnz_tot = sum(pb.nnz for pb in pbs)
    say(f"setup {time.perf_counter() - t0:.1f}s: {len(pbs)} problems, sides {sides.min()}..{sides.max()}, "
        f"{n_tot} DOFs, {nnz_tot} nnz")
    batch = rh.Batch(ctx, mats)
    db = ctx.from_numpy(batch.pack(bs))
    dx = ctx.from_numpy(np.zeros(batch.total_rows))
    batch.cg(db, dx, tol=1e-8, max_iters=5000)  # warm-up (graph capture, workspaces)
    dx.upload(np.zeros(batch.total_rows))
    ctx.sync()
    tb = time.perf_counter()
    reps = batch.cg(db, dx, tol=1e-8, max_iters=5000)
    tb = time.perf_counter() - tb
    xs = batch.unpack(dx.numpy())
    ts, iters_b, iters_s = 0.0, 0, 0
    for i, (pb, K, b) in enumerate(zip(pbs, mats, bs)):
        x1 = pb.vector(0.0)
        d_b = ctx.from_numpy(b)
        ctx.sync()
        t1 = time.perf_counter()
        r1 = K.cg(d_b, x1, tol=1e-8, max_iters=5000)
        ts += time.perf_counter() - t1
        assert reps[i].converged == 1 and r1.converged == 1
        assert reps[i].iterations == r1.iterations, (i, reps[i].iterations, r1.iterations)
        assert np.array_equal(xs[i], x1.numpy()), i
        iters_b += reps[i].iterations
        iters_s += r1.iterations
        x1.free()
        d_b.free()
    say(f"batched: {tb*1e3:.0f} ms for {iters_b} problem-iterations "
        f"({sum(pb.n_free * r.iterations for pb, r in zip(pbs, reps)) / tb:.3e} DOF-iters/s); "
        f"256 standalone solves: {ts*1e3:.0f} ms; speed-up {ts / tb:.2f}x")
    # FP32 bulk with FP64 refinement: the batched mixed solve, every problem
    # bitwise its standalone mixed solve, within 1e-6 of the FP64 solution
    dm = ctx.from_numpy(np.zeros(batch.total_rows))
    tm = time.perf_counter()
    mreps, mcyc = batch.cg_mixed(db, dm, tol=1e-8, max_iters=5000)
    tm = time.perf_counter() - tm
    xm = batch.unpack(dm.numpy())
    for i, (pb, K, b) in enumerate(zip(pbs, mats, bs)):
        assert mreps[i].converged == 1
        assert np.max(np.abs(xm[i] - xs[i])) <= 1e-6 * np.max(np.abs(xs[i])), i
        if i % 16 == 0:
            x1 = pb.vector(0.0)
            d_b = ctx.from_numpy(b)
            r1, c1 = K.cg_mixed(d_b, x1, tol=1e-8, max_iters=5000)
            assert mreps[i].iterations == r1.iterations and mcyc[i] == c1, i
            assert np.array_equal(xm[i], x1.numpy()), i
            x1.free()
            d_b.free()
    say(f"batched mixed: {tm*1e3:.0f} ms (first call, FP32 copies included) for "
        f"{sum(r.iterations for r in mreps)} inner iterations, outer cycles {min(mcyc)}..{max(mcyc)}")


@pytest.mark.gpu
def test_config1_steady_square_64(ctx):
    """The plain FP64 workload as stated: 64x64 Q1 quads on the unit square
    (4,225 nodes, 3,969 free), 2x2 Gauss, k = 1, Dirichlet T = (1-x)(1-y)
    on all four sides, no source, CG + Jacobi to 1e-10 relative from 0.
    The solve is bitwise the oracle's in the kernels' order (iterations,
    report, solution); the discrete solution is the bilinear field itself,
    so the nodal values match (1-x)(1-y) to the solver tolerance; against
    the reference-order oracle at 1e-13 the solutions agree to 1e-12."""
    import residuum_hip as rh
    dim, n, p, nq, tags, g = 2, (64, 64), (1, 1), 2, [0, 1, 2, 3], ["(1-x)*(1-y)"] * 4
    pb = rh.Problem(ctx, dim, n, p, nq, tags, g)
    oc = O.Case(dim, n, p, nq, tags, g, order=O.ORDER_GPU)
    assert pb.n_nodes == 4225 and pb.n_free == 3969
    forms = [(rh.DIFFUSION, 1.0)]
    K, F, U = pb.matrix(), pb.vector(), pb.vector(0.0)
    pb.assemble_matrix(K, forms)
    pb.assemble_load(F)
    pb.apply_essential(F, forms)
    t0 = time.perf_counter()
    rep = K.cg(F, U, tol=1e-10, max_iters=2000)
    dt = time.perf_counter() - t0
    Uo, ro = oc.steady(solver=0, tol=1e-10, max_iters=2000)
    say(f"config 1: {rep.iterations} CG iterations ({ro.iterations} oracle) in {dt*1e3:.1f} ms, "
        f"{pb.n_free * rep.iterations / dt:.3e} DOF-iters/s (one solve, host timing)")
    assert rep.converged == 1 and rep.iterations == ro.iterations
    assert np.array_equal(U.numpy(), Uo)
    xyz = oc.xyz.reshape(-1, 2)[oc.a2t]
    exact = (1 - xyz[:, 0]) * (1 - xyz[:, 1])
    err = np.max(np.abs(U.numpy() - exact))
    say(f"max |T - (1-x)(1-y)| at the nodes: {err:.2e}")
    assert err < 1e-7
    oc.set_order(O.ORDER_REF)
    Ur, _ = oc.steady(solver=0, tol=1e-13, max_iters=4000)
    U2 = pb.vector(0.0)
    K.cg(F, U2, tol=1e-13, max_iters=4000)
    assert np.max(np.abs(U2.numpy() - Ur)) <= 1e-12 * np.max(np.abs(Ur))
    assert np.max(np.abs(U2.numpy() - exact)) < 1e-10


MIXED_CG_STALLS = ("mixed CG on the full FP32 copy of this anisotropic, non-symmetric Jacobian stalls: "
                   "40 refinements without reaching the inner tolerance, Newton rc -2 at 128^2 and 512^2")


@pytest.mark.gpu
@pytest.mark.parametrize("n,lin_solver,lin_tol,ortho", [
    (512, 1, 1e-3, 0), (512, 1, 1e-3, 1), (512, 2, 1e-2, 0), (512, 2, 1e-2, 1), (128, 1, 1e-12, 0),
    (128, 1, 1e-12, 1), (128, 2, 1e-10, 0), (128, 2, 1e-10, 1),
    pytest.param(512, 3, 1e-10, 0, marks=pytest.mark.skip(reason=MIXED_CG_STALLS)),
    pytest.param(128, 3, 1e-10, 0, marks=pytest.mark.skip(reason=MIXED_CG_STALLS))],
    ids=["512-gmres30", "512-gmres30-cgs2", "512-gmres30-mixed", "512-gmres30-mixed-cgs2", "128-gmres30",
         "128-gmres30-cgs2", "128-gmres30-mixed", "128-gmres30-mixed-cgs2", "512-cg-mixed", "128-cg-mixed"])
def test_newton_q2(ctx, n, lin_solver, lin_tol, ortho):
    """Q2 quads on the unit square (3x3 Gauss), k(T) = (1 + 0.5 T)
    diag(1, 0.1), f = 10 sin(pi x) sin(pi y), T = 0 on every side; Newton
    (relative and absolute 1e-10) with GMRES(30) + Jacobi inner solves,
    FP64 (lin_solver 1) or mixed precision (2: FP32 inner GMRES(30), FP64
    refinement).  Checked: convergence, the residual F - K(U) U recomputed by
    the vector assembly, and the two mirror symmetries of the problem.

    Inner tolerance: at 512^2 (1,046,529 DOFs) the solves are inexact
    Newton steps.  Restarted GMRES(30) needs 2.7-4.4x more Arnoldi steps per
    mesh halving on these Jacobians (the CPU growth test) and ~30k steps
    for 1e-3 at 512^2, so FP64 runs to 1e-3 (4 Newton iterations, ~120k
    Arnoldi steps) and the mixed solve to 1e-2 (its inner FP32 solve aims at
    half the outer tolerance, one refinement cycle per Newton step).  At
    128^2 FP64 GMRES(30) runs to 1e-12 and the mixed solve to 1e-10: the
    refinement's attainable relative residual on these Jacobians is about
    1e-12.  The cgs2 cases run the GMRES (FP64, or the mixed solver's FP32
    inner one) with classical Gram-Schmidt reorthogonalised once
    (RSD_ORTHO_CGS2) under the same checks."""
    import residuum_hip as rh
    t0 = time.perf_counter()
    pb = rh.Problem(ctx, 2, (n, n), (2, 2), 3, [0, 1, 2, 3], ["0"] * 4)
    pb.set_conductivity(1.0, k_expr="1.0 + 0.5*T", dk_expr="0.5", tensor=(1.0, 0.0, 0.0, 0.1))
    say(f"setup {time.perf_counter() - t0:.2f}s n_free={pb.n_free} nnz={pb.nnz}")
    src = "10*sin(pi*x)*sin(pi*y)"
    U = pb.vector(0.0)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.set_gmres_orthogonalization(ortho)
    try:
        its, res, lin = pb.newton(U, src=src, abs_tol=1e-10, rel_tol=1e-10, lin_solver=lin_solver, lin_tol=lin_tol,
                                  lin_max=100000, krylov=30, max_newton=30)
    finally:
        ctx.set_gmres_orthogonalization(rh.ORTHO_MGS)
    t1 = time.perf_counter() - t1
    say(f"newton: {its} iterations in {t1*1e3:.0f} ms, inner iterations {list(lin)}, residuals {list(res)}")
    assert its > 0 and (res[-1] < 1e-10 or res[-1] < 1e-10 * res[0])
    # residual recomputed: F - K(U) U (the boundary values are 0, so no lift)
    F, KU = pb.vector(), pb.vector()
    pb.assemble_load(F, src=src)
    pb.assemble_vector(KU, [(rh.DIFFUSION, 1.0)], U=U)
    r = F.numpy() - KU.numpy()
    say(f"recomputed ||F - K(U)U|| = {np.linalg.norm(r):.3e} (first residual {res[0]:.3e})")
    assert np.linalg.norm(r) <= 2e-10 * max(1.0, res[0])
    # mirror symmetries x -> 1 - x and y -> 1 - y of the nodal field
    t2a, a2t, _ = pb.dofs()
    m = 2 * n + 1
    T = np.zeros(m * m)
    T[a2t] = U.numpy()
    T = T.reshape(m, m)
    scale = np.abs(T).max()
    sx = np.abs(T - T[:, ::-1]).max() / scale
    sy = np.abs(T - T[::-1, :]).max() / scale
    say(f"max T {scale:.6f}; symmetry defects x {sx:.2e} y {sy:.2e}")
    assert sx < 1e-8 and sy < 1e-8
