"""Benchmark: CG DOF-iterations/s and SpMV HBM GB/s on MI355X, mixed
precision vs FP64, relative residual 1e-8.

Workload: 3D transient heat on a structured 128^3 Q1 hex mesh of the unit
cube (2,146,689 nodes, 2,113,407 free DOFs), 2x2x2 Gauss, per-element
conductivity lognormal(0, 0.5) from Philox seed 20261201, rho*c = 1, T0 = 0,
T = 1 on x = 0 and T = 0 on x = 1, insulated elsewhere, Backward Euler with
dt = 1e-3, CG + Jacobi to relative residual 1e-8 warm-started each step.
A "step" is one Backward Euler step (assembly of K + M/dt, right-hand side,
essential lift, CG solve), run once in FP64 and once in mixed precision.

Prints one JSON line.  value = mixed-precision CG DOF-iterations per second
(free DOFs x CG iterations / GPU solve time); the FP64 figure, the
whole-step rates and the SpMV roofline sit beside it.
"""
import argparse
import ctypes as C
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "residuum-solver_amd"))

PEAK_HBM_GBS = 8000.0


def hip_events():
    """hipEvent helpers on an explicit stream (the rsd context's stream)."""
    hip = C.CDLL("libamdhip64.so")
    hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
    hip.hipEventSynchronize.argtypes = [C.c_void_p]
    hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]

    def new():
        e = C.c_void_p()
        assert hip.hipEventCreate(C.byref(e)) == 0
        return e

    def rec(e, stream):
        assert hip.hipEventRecord(e, stream) == 0

    def ms(a, b):
        assert hip.hipEventSynchronize(b) == 0
        t = C.c_float()
        assert hip.hipEventElapsedTime(C.byref(t), a, b) == 0
        return float(t.value)
    return new, rec, ms


def hip_graph_capture():
    """capture(fn, reps, stream): reps calls of fn recorded on `stream` into
    an instantiated hipGraph, so that their replay runs the kernels back to
    back with no host call in between (a Python loop of launches cannot keep
    up with a 20 us kernel: events around it would time the host)."""
    hip = C.CDLL("libamdhip64.so")
    hip.hipStreamBeginCapture.argtypes = [C.c_void_p, C.c_int]
    hip.hipStreamEndCapture.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
    hip.hipGraphInstantiate.argtypes = [C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    hip.hipGraphLaunch.argtypes = [C.c_void_p, C.c_void_p]

    def capture(fn, reps, stream):
        assert hip.hipStreamBeginCapture(stream, 1) == 0  # hipStreamCaptureModeThreadLocal
        try:
            for _ in range(reps):
                fn()
        finally:
            g = C.c_void_p()
            rc = hip.hipStreamEndCapture(stream, C.byref(g))
        assert rc == 0
        ge = C.c_void_p()
        assert hip.hipGraphInstantiate(C.byref(ge), g, None, None, 0) == 0

        def replay():
            if hip.hipGraphLaunch(ge, stream) != 0:
                raise RuntimeError("hipGraphLaunch failed")
        return replay
    return capture


def cpu_info():
    model = platform.processor() or "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def _cores():
    """The cores this process may run on (the box's share, not the machine's)."""
    try:
        return sorted(os.sched_getaffinity(0))
    except AttributeError:
        return list(range(os.cpu_count() or 1))


class _pinned:
    """Run a block on one core (the reference is single-threaded): the
    same effect as `taskset -c <core>` for this thread's timing."""

    def __init__(self, core):
        self.core = core

    def __enter__(self):
        self.old = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {self.core})
        return self

    def __exit__(self, *exc):
        os.sched_setaffinity(0, self.old)


def cpu_baseline(pb, K, b_host, iters, kappa, dt):
    """Serial CPU CG on the same assembled matrix, one pinned core, as a
    differential sample: the rate is nfree * (k2 - k1) / (t(k2) - t(k1)),
    so copies and first touches that do not scale with the iterations drop
    out.  The reference's own cg::Solver with Jacobi (oracle/_ref/libref.so,
    compiled from the reference's headers at -O3) when that harness is
    built, kind "reference"; the oracle's C restatement beside it (the port
    over reference time ratio is the fidelity of the port); the oracle
    alone, kind "port", when the harness is absent.  Also the reference's
    assembleMatrix (diffusion with the per-element conductivity) over every
    element, in elements/s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O
    core = _cores()[0]
    rp, ci = pb.pattern()
    val = K.values()
    n = pb.n_free
    k1 = 20

    def port(k):
        x = np.zeros(n)
        rep = O.OrcReport()
        t0 = time.perf_counter()
        O.lib().orc_cg(C.c_int64(n), O.lp(rp), O.lp(ci), O.dp(val), O.dp(b_host), O.dp(x), C.c_double(1e-300),
                       C.c_int64(k), 1, O.ORDER_REF, C.byref(rep), None)
        return time.perf_counter() - t0, int(rep.iterations)

    out = {}
    with _pinned(core):
        if O.have_ref():
            dim_n = round(pb.n_elem ** (1.0 / 3.0))
            r = O.RefCase(3, (dim_n, dim_n, dim_n), (1, 1, 1), 2, [0, 1], [6, 0])
            assert r.n_free == n and r.nnz == pb.nnz, "reference harness built a different problem"

            def ref(k):
                t0 = time.perf_counter()
                _, _, it = r.solve(0, 1, val, b_host, tol=1e-300, max_iters=k)
                return time.perf_counter() - t0, it
            (ta, ia), (tb, ib) = ref(k1), ref(iters)
            out["ref_rate"], out["ref_iters"], out["ref_s"] = n * (ib - ia) / (tb - ta), ib - ia, tb - ta
            # the Backward Euler operator K + M/dt (the harness's set 2),
            # the matrix each step assembles; its values against the
            # device's (reference vs GPU summation order)
            t0 = time.perf_counter()
            vref = r.assemble_matrix(2, 1.0 / dt, kappa=kappa)
            out["ref_asm_s"] = time.perf_counter() - t0
            out["ref_asm_elements_per_s"] = pb.n_elem / out["ref_asm_s"]
            out["ref_vs_gpu_matrix_max_rel_diff"] = float(np.max(np.abs(vref - val)) / np.max(np.abs(vref)))
            del r
        kp = iters if "ref_rate" not in out else max(k1 + 100, iters // 2)
        (ta, ia), (tb, ib) = port(k1), port(kp)
        out["port_rate"], out["port_iters"], out["port_s"] = n * (ib - ia) / (tb - ta), ib - ia, tb - ta
    out["core"] = core
    return out


def oracle_be_step(n, path):
    """Child process (started before the parent touches the GPU, pinned to
    its own core): one Backward Euler step of the benchmark problem through
    the oracle in the GPU summation order, U saved to `path` (.npy).  The
    parent compares its FP64 step (bitwise expected) and its mixed step."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O
    kappa = np.exp(0.5 * O.normal_field(20261201, 0, n ** 3))
    c = O.Case(3, (n, n, n), (1, 1, 1), 2, [0, 1], ["1", "0"], kappa=kappa, order=O.ORDER_GPU)
    out = c.be_run(0.0, 1e-3, 1e-3, cg_tol=1e-8, cg_max=5000)
    np.save(path, out["U"])
    np.save(path[:-4] + "_iters.npy", out["iters"])


def start_oracle_check(n):
    import tempfile
    cores = _cores()
    d = tempfile.mkdtemp(prefix="rsd_bench_")
    path = os.path.join(d, "oracle_U.npy")
    proc = subprocess.Popen([sys.executable, os.path.abspath(__file__), "--oracle-be-step", str(n), path],
                            preexec_fn=(lambda: os.sched_setaffinity(0, {cores[-1]})) if len(cores) > 1 else None)
    return proc, path, d


def finish_oracle_check(proc, path, d, U64, Umx, it64, itmx, timeout=600):
    import shutil
    try:
        rc = proc.wait(timeout=timeout)
        if rc != 0:
            return {"error": "oracle child exited with %d" % rc}
        Uo = np.load(path)
        ito = np.load(path[:-4] + "_iters.npy")
        scale = max(float(np.max(np.abs(Uo))), 1e-300)
        return {"steps": 1, "oracle": "C restatement, GPU summation order, FP64, 1 core",
                "oracle_cg_iters": int(ito[0]), "fp64_cg_iters": int(it64), "mixed_inner_iters": int(itmx),
                "fp64_bitwise": bool(np.array_equal(U64, Uo)),
                "fp64_max_rel_err": float(np.max(np.abs(U64 - Uo)) / scale),
                "mixed_max_rel_err": float(np.max(np.abs(Umx - Uo)) / scale)}
    except subprocess.TimeoutExpired:
        proc.kill()
        return {"error": "oracle child timed out"}
    finally:
        shutil.rmtree(d, ignore_errors=True)


LAYOUTS = ["int32 columns", "int16 offsets", "shared offset patterns", "symmetric half storage",
           "symmetric half storage, uniform 27-point stencil"]


def measured_traffic(n, kernel):
    """HBM bytes per launch of `kernel` at this size, from the committed
    FETCH_SIZE / WRITE_SIZE counter profile of the benchmark (written by
    --summarize-rocprof hbm, scaled by the calibration copy's counters);
    None when no profile of this size records that kernel."""
    path = os.path.join(ROOT, "profiles", "hbm_counters_be_%dcube.json" % n)
    try:
        with open(path) as f:
            prof = json.load(f)
        if prof.get("kernel") != kernel:
            return None
        return float(prof["traffic_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def run_batched(args, rh, ctx, stream, new_ev, rec_ev, ev_ms):
    """The batched workload (not the default line): 256 steady heat problems
    on Q1 hex unit cubes, sides 16..64 (Philox, seed 20261202), per-element
    conductivity lognormal(0, 1), nodal source U[0, 1], T = 0 on every face,
    CG + Jacobi to 1e-8 from x = 0; one batched solve under the autotuned
    schedule against the same problems solved one after another, then the
    batched mixed-precision solve (FP32 bulk, FP64 refinement) against the
    same FP64 solves."""
    sides = (16 + np.floor(rh.philox_uniform(20261202, 0, args.batch_problems) * 49)).astype(int)
    t0 = time.perf_counter()
    pbs, mats, bs = [], [], []
    for i, s_ in enumerate(sides):
        s_ = int(s_)
        pb = rh.Problem(ctx, 3, (s_, s_, s_), (1, 1, 1), 2, list(range(6)), ["0"] * 6)
        pb.set_conductivity(1.0, kappa=np.exp(rh.philox_normal(20261202, 1 + i, pb.n_elem)))
        K = pb.matrix()
        pb.assemble_matrix(K, [(rh.DIFFUSION, 1.0)])
        F = pb.vector()
        pb.assemble_load(F, nodal=ctx.from_numpy(rh.philox_uniform(20261202, 1000 + i, pb.n_nodes)))
        pbs.append(pb)
        mats.append(K)
        bs.append(F)
    ctx.sync()
    setup_s = time.perf_counter() - t0
    rows = [pb.n_free for pb in pbs]
    nnz = sum(pb.nnz for pb in pbs)
    batch = rh.Batch(ctx, mats)
    sched, tune_us, hit = batch.autotune(cache=args.tune_cache, iters=32)
    db = ctx.from_numpy(batch.pack([F.numpy() for F in bs]))
    dx = ctx.from_numpy(np.zeros(batch.total_rows))
    batch.cg(db, dx, tol=1e-8, max_iters=5000)  # warm-up
    dx = ctx.from_numpy(np.zeros(batch.total_rows))
    ctx.sync()
    a, b = new_ev(), new_ev()
    rec_ev(a, stream)
    reps = batch.cg(db, dx, tol=1e-8, max_iters=5000)
    rec_ev(b, stream)
    batched_ms = ev_ms(a, b)
    its = [int(r.iterations) for r in reps]
    assert all(r.converged for r in reps)
    xs = batch.unpack(dx.numpy())
    # the same problems one after another
    xs1 = [pb.vector(0.0) for pb in pbs]
    ctx.sync()
    rec_ev(a, stream)
    its1 = [int(K.cg(F, x1, tol=1e-8, max_iters=5000).iterations) for K, F, x1 in zip(mats, bs, xs1)]
    rec_ev(b, stream)
    single_ms = ev_ms(a, b)
    same_bits = all(np.array_equal(x, x1.numpy()) for x, x1 in zip(xs, xs1)) and its == its1
    dof_iters = float(sum(n_ * k for n_, k in zip(rows, its)))
    print(f"# batched: {len(pbs)} problems, {sum(rows)} rows, {nnz} nnz, setup {setup_s:.1f} s; schedule {sched} "
          f"({'cached' if hit else 'tuned'}, {tune_us:.0f} us/iter); batched {batched_ms:.1f} ms, one by one "
          f"{single_ms:.1f} ms; bitwise equal: {same_bits}", file=sys.stderr, flush=True)
    # FP32 bulk with FP64 refinement: the batched mixed solve (first call
    # converts the FP32 copies and sets up the tables, so it is a warm-up)
    dm = ctx.from_numpy(np.zeros(batch.total_rows))
    batch.cg_mixed(db, dm, tol=1e-8, max_iters=5000)
    dm = ctx.from_numpy(np.zeros(batch.total_rows))
    ctx.sync()
    rec_ev(a, stream)
    mreps, mcyc = batch.cg_mixed(db, dm, tol=1e-8, max_iters=5000)
    rec_ev(b, stream)
    mixed_ms = ev_ms(a, b)
    assert all(r.converged for r in mreps)
    mits = [int(r.iterations) for r in mreps]
    xm = batch.unpack(dm.numpy())
    m_same = True
    for K, F, x, it_, c_ in zip(mats, bs, xm, mits, mcyc):
        x1 = K.pb.vector(0.0)
        r1, c1 = K.cg_mixed(F, x1, tol=1e-8, max_iters=5000)
        m_same = m_same and np.array_equal(x, x1.numpy()) and int(r1.iterations) == it_ and c1 == c_
    m_diff = max(float(np.max(np.abs(x - x64)) / max(np.max(np.abs(x64)), 1e-300)) for x, x64 in zip(xm, xs))
    m_dof_iters = float(sum(n_ * k for n_, k in zip(rows, mits)))
    print(f"# batched mixed: {mixed_ms:.1f} ms ({sum(mits)} inner iterations, {max(mits)} at most, outer cycles "
          f"{min(mcyc)}..{max(mcyc)}); bitwise equal to the standalone mixed solves: {m_same}; max relative "
          f"difference to FP64 {m_diff:.2e}", file=sys.stderr, flush=True)
    line = {
        "metric": "batched CG DOF-iters/s (256 uneven Q1 hex problems, CG+Jacobi to 1e-8)",
        "value": dof_iters / (batched_ms * 1e-3),
        "unit": "DOF-iters/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic (seeded sides, lognormal(0,1) conductivity, U[0,1] nodal source)",
        "config": {"workload": "batched steady heat, %d Q1 hex unit cubes, sides 16..64" % len(pbs),
                   "rows": int(sum(rows)), "nnz": int(nnz)},
        "batched_ms": batched_ms, "one_by_one_ms": single_ms, "speedup": single_ms / batched_ms,
        "iterations": {"sum": int(sum(its)), "max": int(max(its))},
        "schedule": {"chunk": sched[0], "unroll": sched[1], "order": sched[2], "xcd": sched[3],
                     "from_cache": hit, "tuned_us_per_iter": tune_us},
        "bitwise_equal_to_one_by_one": bool(same_bits),
        "mixed": {"batched_ms": mixed_ms, "dof_iters_per_s": m_dof_iters / (mixed_ms * 1e-3),
                  "iterations": {"sum": int(sum(mits)), "max": int(max(mits))},
                  "outer_cycles": {"min": int(min(mcyc)), "max": int(max(mcyc))},
                  "vs_one_by_one_fp64": single_ms / mixed_ms, "vs_batched_fp64": batched_ms / mixed_ms,
                  "bitwise_equal_to_standalone_mixed": bool(m_same), "max_rel_diff_vs_fp64": m_diff},
    }
    print(json.dumps(line), flush=True)


def run_newton(args, rh, ctx, stream, new_ev, rec_ev, ev_ms):
    """The nonlinear workload (not the default line): 512x512 Q2 quads on the
    unit square (3x3 Gauss), k(T) = (1 + 0.5 T) diag(1, 0.1), source
    10 sin(pi x) sin(pi y), T = 0 on the boundary; Newton to relative and
    absolute 1e-10 with GMRES(30) + Jacobi inner solves: FP64 to 1e-3
    (inexact Newton steps: GMRES(30) stagnates near 1e-4 at this size) and
    mixed precision (FP32 inner GMRES, FP64 refinement) to 1e-2, each with the
    reference's modified Gram-Schmidt and with classical Gram-Schmidt
    reorthogonalised once."""
    n = args.newton_n
    src = "10*sin(pi*x)*sin(pi*y)"
    pb = rh.Problem(ctx, 2, (n, n), (2, 2), 3, [0, 1, 2, 3], ["0"] * 4)
    pb.set_conductivity(1.0, k_expr="1.0 + 0.5*T", dk_expr="0.5", tensor=(1.0, 0.0, 0.0, 0.1))
    runs = {}
    for name, ortho, lin_solver, lin_tol in (("mgs", rh.ORTHO_MGS, 1, 1e-3), ("cgs2", rh.ORTHO_CGS2, 1, 1e-3),
                                             ("mixed_mgs", rh.ORTHO_MGS, 2, 1e-2),
                                             ("mixed_cgs2", rh.ORTHO_CGS2, 2, 1e-2)):
        U = pb.vector(0.0)
        ctx.set_gmres_orthogonalization(ortho)
        ctx.sync()
        e0, e1 = new_ev(), new_ev()
        rec_ev(e0, stream)
        its, res, lin = pb.newton(U, src=src, abs_tol=1e-10, rel_tol=1e-10, lin_solver=lin_solver, lin_tol=lin_tol,
                                  lin_max=100000, krylov=30, max_newton=30)
        rec_ev(e1, stream)
        ms = ev_ms(e0, e1)
        ctx.set_gmres_orthogonalization(rh.ORTHO_MGS)
        steps = int(np.sum(lin))
        F, KU = pb.vector(), pb.vector()
        pb.assemble_load(F, src=src)
        pb.assemble_vector(KU, [(rh.DIFFUSION, 1.0)], U=U)
        rr = float(np.linalg.norm(F.numpy() - KU.numpy()))
        runs[name] = dict(ms=ms, newton_iterations=int(its), arnoldi_steps=steps, us_per_arnoldi_step=ms * 1e3 / steps,
                          residuals=[float(r) for r in res], recomputed_residual=rr, U=U.numpy())
        print(f"# newton {name}: {its} iterations, {steps} Arnoldi steps in {ms:.0f} ms "
              f"({ms * 1e3 / steps:.1f} us per step), recomputed ||F - K(U)U|| {rr:.2e}", file=sys.stderr, flush=True)
    a, b = runs["mgs"].pop("U"), runs["cgs2"].pop("U")
    am, bm = runs["mixed_mgs"].pop("U"), runs["mixed_cgs2"].pop("U")
    line = {
        "metric": "Newton + GMRES(30) wall time (512^2 Q2 nonlinear steady heat)", "value": runs["cgs2"]["ms"],
        "unit": "ms", "n_gpus": 1, "higher_is_better": False, "dtype": "f64",
        "data": "synthetic (the nonlinear configuration's closed-form coefficients)",
        "config": {"workload": f"nonlinear steady heat, {n}x{n} Q2 quads, k(T) anisotropic, Newton 1e-10, "
                               "GMRES(30)+Jacobi inner 1e-3", "dofs": pb.n_free, "nnz": pb.nnz},
        "mgs": runs["mgs"], "cgs2": runs["cgs2"], "mgs_over_cgs2_time": runs["mgs"]["ms"] / runs["cgs2"]["ms"],
        "cgs2_vs_mgs_solution_max_rel_diff": float(np.max(np.abs(a - b)) / max(np.max(np.abs(a)), 1e-300)),
        "mixed_mgs": runs["mixed_mgs"], "mixed_cgs2": runs["mixed_cgs2"],
        "mixed_cgs2_vs_fp64_mgs_solution_max_rel_diff": float(np.max(np.abs(bm - a)) / max(np.max(np.abs(a)), 1e-300)),
    }
    print(json.dumps(line), flush=True)


def run_adjoint(args, rh, ctx, stream, new_ev, rec_ev, ev_ms):
    """The adjoint-gradient workload (not the default line): 96^3 Q1 hex on
    the unit cube, kappa_e = exp(theta_e) with theta a Gaussian random field
    (std 0.5, correlation length 0.1), rho c = 1, Backward Euler dt = 5e-3
    for 20 steps from T = 0 with T = 1 on z = 0, loss 1/2 ||T_N[obs] -
    T_obs||^2 over 4096 seeded free nodes, T_obs from the forward run with a
    perturbed field; solves to 1e-13 relative (the gradient checks need
    them converged).  Times the forward loss and the forward plus adjoint
    sweep with HIP events, and checks the gradient against central
    differences (h = 1e-6) along a few of the seeded directions the
    full-size test uses."""
    from residuum_hip.fields import gaussian_random_field, sample_without_replacement
    n = args.adjoint_n
    dt, steps, tol = 5e-3, 20, 1e-13
    pb = rh.Problem(ctx, 3, (n, n, n), (1, 1, 1), 2, [4], ["1"])
    pb.set_mass(1.0, 1.0)
    theta = gaussian_random_field(20261203, (n, n, n), 0.1, 0.5)
    xi = gaussian_random_field(20261205, (n, n, n), 0.1, 0.5)
    obs = sample_without_replacement(20261204, pb.n_free, 4096)
    U0 = pb.vector(0.0)
    dUN = ctx.empty(pb.n_free)
    dk = ctx.from_numpy(np.exp(theta + 0.1 * xi))
    pb.adjoint(dk, U0, 0.0, dt, steps, np.zeros(0, np.int64), np.zeros(0), tol=tol, want_grad=False, d_UN=dUN)
    t_obs = dUN.numpy()[obs]
    dg = ctx.empty(pb.n_elem)

    def timed(th, grad):
        dk.upload(np.exp(th))
        ctx.sync()
        e0, e1 = new_ev(), new_ev()
        rec_ev(e0, stream)
        L = pb.adjoint(dk, U0, 0.0, dt, steps, obs, t_obs, tol=tol, want_grad=grad, d_grad=dg if grad else None)
        rec_ev(e1, stream)
        return L, ev_ms(e0, e1)

    timed(theta, True)  # warm-up
    reps = max(1, args.adjoint_reps)
    grad_ms, fwd_ms = [], []
    for _ in range(reps):
        L, ms = timed(theta, True)
        grad_ms.append(ms)
    g = dg.numpy()
    for _ in range(reps):
        L2, ms = timed(theta, False)
        fwd_ms.append(ms)
    print(f"# adjoint {n}^3: loss {L:.6e}; forward {np.median(fwd_ms):.1f} ms, forward + adjoint sweep "
          f"{np.median(grad_ms):.1f} ms", file=sys.stderr, flush=True)
    h = 1e-6
    checks = []
    for j in range(args.adjoint_directions):
        d = rh.philox_normal(20261206, j, pb.n_elem)
        fd = (timed(theta + h * d, False)[0] - timed(theta - h * d, False)[0]) / (2 * h)
        gd = float(g @ d)
        checks.append(dict(direction=j, adjoint=gd, central_fd=fd, rel_diff=abs(fd - gd) / abs(gd)))
        print(f"# direction {j}: adjoint {gd:+.10e}  central FD {fd:+.10e}  rel diff {checks[-1]['rel_diff']:.2e}",
              file=sys.stderr, flush=True)
    gms = float(np.median(grad_ms))
    line = {
        "metric": "adjoint gradient wall time (forward with checkpoints + adjoint sweep)", "value": gms,
        "unit": "ms", "n_gpus": 1, "higher_is_better": False, "dtype": "f64",
        "data": "synthetic (seeded Gaussian random log-conductivity, seeded observations)",
        "config": {"workload": f"adjoint gradient, {n}^3 Q1 hex, Backward Euler dt=5e-3 x {steps}, "
                               "4096 observations, solves to 1e-13", "dofs": pb.n_free, "elements": pb.n_elem},
        "loss": L, "forward_loss_repeats_bitwise": bool(L2 == L), "forward_ms": float(np.median(fwd_ms)),
        "gradient_ms": gms, "gradient_over_forward": gms / float(np.median(fwd_ms)),
        "element_steps_per_s": pb.n_elem * steps * 2 / (gms * 1e-3),
        "fd_check": checks,
    }
    print(json.dumps(line), flush=True)


def renumbered_workload(rh, ctx, stream, n, kappa_elem, dt, capture, new_ev, rec_ev, ev_ms, reps, bytes_of):
    """The benchmark's operator on the same 128^3 mesh with its nodes
    renumbered in 8^3-node tiles (tile-major, lexicographic inside a tile):
    the locality of a good unstructured ordering, but the rows no longer
    share a few hundred column-offset patterns, so the sliced ELL falls back
    to explicit columns and the FP32 copy to the full layout.  Times the
    FP64 and FP32 products and one step's FP64 and mixed solves."""
    xyz, ien, rng = rh.block_mesh(3, (n, n, n), (0.0, 1.0, 0.0, 1.0, 0.0, 1.0), (1, 1, 1))
    m = n + 1
    idx = np.arange(m ** 3)
    i, j, k = idx % m, (idx // m) % m, idx // (m * m)
    t = 8
    nt = (m + t - 1) // t
    tile = ((k // t) * nt + (j // t)) * nt + (i // t)
    inner = ((k % t) * t + (j % t)) * t + (i % t)
    order = np.lexsort((inner, tile))       # old node ids in the new order
    new_of = np.empty_like(order)
    new_of[order] = np.arange(order.size)
    xyz2 = xyz.reshape(-1, 3)[order].ravel()
    pb = rh.Problem.from_mesh(ctx, 3, (1, 1, 1), 2, xyz2, new_of[ien], rng, [0, 1], ["1", "0"])
    pb.set_conductivity(1.0, kappa=kappa_elem)
    pb.set_mass(1.0, 1.0)
    nfree, nnz = pb.n_free, pb.nnz
    col_mode, _, _, _ = pb.layout()
    mode32, _, _, _ = pb.layout_f32()
    K = pb.matrix()
    pb.assemble_matrix(K, [(rh.DIFFUSION, 1.0), (rh.MASS | rh.SCALED, 1.0 / dt)], time=dt)
    K.to_f32()
    L = rh.lib()
    x64, y64 = pb.vector(1.0), pb.vector()
    x32, y32 = ctx.empty(nfree, np.float32), ctx.empty(nfree, np.float32)
    ctx.fill(x32, 1.0)
    out = {"numbering": "8^3-node tiles", "layout_f64": LAYOUTS[col_mode], "layout_f32": LAYOUTS[mode32]}
    for name, fn, byts in (("spmv_f64", lambda: rh.check(L.rsd_spmv_f64(ctx.h, K.h, rh.dptr(x64), rh.dptr(y64))),
                            bytes_of(col_mode, 8)),
                           ("spmv_f32", lambda: rh.check(L.rsd_spmv_f32(ctx.h, K.h, rh.dptr(x32), rh.dptr(y32))),
                            bytes_of(mode32, 4))):
        for _ in range(3):
            fn()
        ctx.sync()
        replay = capture(fn, reps, stream)
        replay()
        ctx.sync()
        a, b = new_ev(), new_ev()
        rec_ev(a, stream)
        replay()
        rec_ev(b, stream)
        us = ev_ms(a, b) / reps * 1e3
        out[name] = {"avg_us": us, "bytes": byts, "gbs": byts / us * 1e-3, "frac": byts / us * 1e-3 / PEAK_HBM_GBS}
    # one Backward Euler step's solves from T = 0, both precisions
    for mixed in (False, True):
        U = pb.vector(0.0)
        steps, iters = pb.be_run(U, 0.0, dt, dt, tol=1e-8, max_iters=5000, mixed=mixed)
        tm = pb.last_timing_ms
        out["mixed" if mixed else "fp64"] = {"iters": int(iters[0]), "solve_ms": tm["solve"],
                                             "dof_iters_per_s": nfree * int(iters[0]) / (tm["solve"] * 1e-3)}
    out["mixed_over_fp64_dof_iters_per_s"] = out["mixed"]["dof_iters_per_s"] / out["fp64"]["dof_iters_per_s"]
    print(f"# renumbered mesh: f64 {LAYOUTS[col_mode]} {out['spmv_f64']['avg_us']:.1f} us "
          f"({out['spmv_f64']['frac']:.2f} of peak), f32 {LAYOUTS[mode32]} {out['spmv_f32']['avg_us']:.1f} us "
          f"({out['spmv_f32']['frac']:.2f}); one step: fp64 {out['fp64']['iters']} iters "
          f"{out['fp64']['solve_ms']:.1f} ms, mixed {out['mixed']['iters']} {out['mixed']['solve_ms']:.1f} ms",
          file=sys.stderr, flush=True)
    K.close()
    pb.close()
    return out


def _short_kernel(name):
    """`void rsd::k_spmv_f64<2>(long, ...)` -> `k_spmv_f64<2>`."""
    name = name.strip().strip('"')
    if name.startswith("void "):
        name = name[5:]
    name = name.replace("rsd::", "")
    depth = 0
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return name[:i].strip()
    return name.strip()


def _csv_rows(dirname, suffix):
    import csv
    import glob
    rows = []
    for path in sorted(glob.glob(os.path.join(dirname, "**", "*" + suffix), recursive=True)):
        with open(path, newline="") as f:
            rows += [{k.strip().lower(): v for k, v in r.items() if k} for r in csv.DictReader(f)]
    return rows


def _col(row, *names, default=None):
    for n in names:
        if n in row and row[n] not in (None, ""):
            return row[n]
    return default


def summarize_rocprof(argv):
    """Profile summaries from rocprofv3 CSV output (--output-format csv):
      trace DIR OUT.md [TITLE [COMMAND]]
          kernel_stats (calls, average / min / max duration, share of GPU
          time) joined with kernel_trace's LDS and scratch sizes per
          workgroup;
      hbm FETCH_DIR WRITE_DIR OUT.md [TITLE [COMMAND [KERNEL=BYTES ...]]]
          FETCH_SIZE and WRITE_SIZE (KiB per dispatch) per kernel and grid,
          scaled by what the same counters read for the calibration copy
          (k_copy over 2^26 doubles: 512 MiB read and 512 MiB written);
          KERNEL=BYTES pairs give algorithmic bytes per launch, and a JSON
          beside OUT.md records the first such kernel's traffic."""
    mode = argv[0]
    if mode == "trace":
        d, out = argv[1], argv[2]
        title = argv[3] if len(argv) > 3 else "Kernel trace summary"
        cmd = argv[4] if len(argv) > 4 else None
        notes = [a[5:] for a in argv[5:] if a.startswith("note:")]
        stats = _csv_rows(d, "kernel_stats.csv")
        trace = _csv_rows(d, "kernel_trace.csv")
        res = {}
        for r in trace:
            k = _short_kernel(_col(r, "kernel_name", default=""))
            lds = int(float(_col(r, "lds_block_size", "group_segment_size", default=0)))
            sc = int(float(_col(r, "scratch_size", "private_segment_size", default=0)))
            res[k] = (lds, sc)
        lines = [f"# {title}", ""]
        if cmd:
            lines += ["Command:", "", "    " + cmd, ""]
        lines += [f"Summarised by `python bench.py --summarize-rocprof trace <dir> <out.md>` from `{os.path.basename(d.rstrip('/'))}`.", ""]
        for t in notes:
            lines += [t, ""]
        lines += [
                  "| kernel | calls | avg us | min us | max us | share of GPU time | LDS B | scratch B |",
                  "|---|---|---|---|---|---|---|---|"]
        stats.sort(key=lambda r: -float(_col(r, "totaldurationns", default=0)))
        for r in stats:
            k = _short_kernel(_col(r, "name", "kernel_name", default=""))
            lds, sc = res.get(k, ("-", "-"))
            lines.append(f"| `{k}` | {int(float(_col(r, 'calls', default=0)))} | "
                         f"{float(_col(r, 'averagens', default=0)) / 1e3:.1f} | "
                         f"{float(_col(r, 'minns', default=0)) / 1e3:.1f} | {float(_col(r, 'maxns', default=0)) / 1e3:.1f} | "
                         f"{float(_col(r, 'percentage', default=0)):.2f}% | {lds} | {sc} |")
        with open(out, "w") as f:
            f.write("\n".join(lines) + "\n")
        print(f"wrote {out} ({len(stats)} kernels)")
        return
    if mode == "hbm":
        fd, wd, out = argv[1], argv[2], argv[3]
        title = argv[4] if len(argv) > 4 else "HBM traffic per launch"
        cmd = argv[5] if len(argv) > 5 else None
        alg, notes = {}, []
        for a in argv[6:]:
            if a.startswith("note:"):
                notes.append(a[5:])
                continue
            k, _, v = a.rpartition("=")
            alg[k] = float(v)

        def per_kernel(d, counter):
            acc = {}
            for r in _csv_rows(d, "counter_collection.csv"):
                if _col(r, "counter_name", default="") != counter:
                    continue
                k = (_short_kernel(_col(r, "kernel_name", default="")), int(float(_col(r, "grid_size", default=0))))
                s_, n_ = acc.get(k, (0.0, 0))
                acc[k] = (s_ + float(_col(r, "counter_value", default=0)), n_ + 1)
            return {k: (v / n, n) for k, (v, n) in acc.items()}
        fetch, write = per_kernel(fd, "FETCH_SIZE"), per_kernel(wd, "WRITE_SIZE")
        calib = ("k_copy", 1 << 26)
        if calib not in fetch or calib not in write:
            raise SystemExit("no calibration copy (k_copy over 2^26 doubles) in the counter output")
        fs, ws = 524288.0 / fetch[calib][0], 524288.0 / write[calib][0]
        lines = [f"# {title}", ""]
        if cmd:
            lines += ["Commands:", ""] + ["    " + c for c in cmd.split(" ;; ")] + [""]
        lines += [f"Summarised by `python bench.py --summarize-rocprof hbm <fetch dir> <write dir> <out.md>`.", "",
                  f"Calibration: `k_copy` of 2^26 doubles (512 MiB read, 512 MiB written) reads FETCH_SIZE = "
                  f"{fetch[calib][0]:.1f} KiB and WRITE_SIZE = {write[calib][0]:.1f} KiB per dispatch, so FETCH_SIZE "
                  f"is scaled by {fs:.4f} and WRITE_SIZE by {ws:.4f}.", ""]
        for t in notes:
            lines += [t, ""]
        lines += ["| kernel (grid) | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes (scaled) | algorithmic bytes | traffic / algorithmic |",
                  "|---|---|---|---|---|---|---|"]
        rows = []
        for k in fetch:
            if k not in write:
                continue
            byts = 1024.0 * (fetch[k][0] * fs + write[k][0] * ws)
            rows.append((byts, k))
        rows.sort(reverse=True)
        record = None
        for byts, k in rows:
            a = alg.get(k[0])
            lines.append(f"| `{k[0]}` ({k[1]}) | {fetch[k][1]} | {fetch[k][0]:.1f} | {write[k][0]:.1f} | "
                         f"{byts / 1e6:.1f} MB | {a / 1e6:.1f} MB | {byts / a:.3f} |" if a else
                         f"| `{k[0]}` ({k[1]}) | {fetch[k][1]} | {fetch[k][0]:.1f} | {write[k][0]:.1f} | "
                         f"{byts / 1e6:.1f} MB | - | - |")
            if a and record is None and k[0] == next(iter(alg)):
                record = {"kernel": k[0], "grid": k[1], "fetch_kib": fetch[k][0], "write_kib": write[k][0],
                          "fetch_scale": fs, "write_scale": ws, "traffic_bytes_per_launch": byts,
                          "algorithmic_bytes_per_launch": a, "source": os.path.basename(out)}
        with open(out, "w") as f:
            f.write("\n".join(lines) + "\n")
        if record is not None:
            with open(out[:-3] + ".json", "w") as f:
                json.dump(record, f, indent=1)
        print(f"wrote {out} ({len(rows)} kernels)")
        return
    raise SystemExit("--summarize-rocprof: mode is 'trace' or 'hbm'")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="timed Backward Euler steps")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128, help="elements per edge")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--cpu-iters", type=int, default=320, help="oracle CG iterations timed for the CPU baseline (about 10 s)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-matrix-free", action="store_true", help="skip the matrix-free operator solve")
    ap.add_argument("--workload", choices=["be128", "batched", "newton", "adjoint"], default="be128",
                    help="be128: the Backward Euler line (default); batched: the batched-solve workload; "
                         "newton: the nonlinear Q2 workload with both GMRES orthogonalisations; "
                         "adjoint: the adjoint-gradient workload")
    ap.add_argument("--newton-n", type=int, default=512, help="elements per edge of the newton workload")
    ap.add_argument("--adjoint-n", type=int, default=96, help="elements per edge of the adjoint workload")
    ap.add_argument("--adjoint-reps", type=int, default=3, help="timed gradients (and forward runs)")
    ap.add_argument("--adjoint-directions", type=int, default=4,
                    help="central-difference directions checked against the gradient")
    ap.add_argument("--batch-problems", type=int, default=256)
    ap.add_argument("--tune-cache", default=None, help="JSON schedule cache for the batched workload")
    ap.add_argument("--no-renumbered", action="store_true",
                    help="skip the same operator on a renumbered mesh (the fallback layouts)")
    ap.add_argument("--no-oracle-check", action="store_true",
                    help="skip the one-step comparison with the oracle (a CPU child process, ~1 min)")
    ap.add_argument("--oracle-be-step", nargs=2, metavar=("N", "OUT"), help=argparse.SUPPRESS)
    ap.add_argument("--summarize-rocprof", nargs="+", metavar="ARG",
                    help="write profile summaries from rocprofv3 CSV output instead of running: "
                         "'trace DIR OUT.md [TITLE]' or 'hbm FETCH_DIR WRITE_DIR OUT.md [TITLE]'")
    args = ap.parse_args()
    if args.oracle_be_step:
        oracle_be_step(int(args.oracle_be_step[0]), args.oracle_be_step[1])
        return
    if args.summarize_rocprof:
        summarize_rocprof(args.summarize_rocprof)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    # the oracle's step runs in a child process beside the GPU work; it is
    # started before this process touches the GPU
    check = None
    if args.workload == "be128" and not args.no_cpu and not args.no_oracle_check:
        check = start_oracle_check(args.n)
    import residuum_hip as rh
    new_ev, rec_ev, ev_ms = hip_events()
    ctx = rh.Ctx(0)
    stream = C.c_void_p()
    rh.check(rh.lib().rsd_ctx_stream(ctx.h, C.byref(stream)))
    if args.workload == "batched":
        run_batched(args, rh, ctx, stream, new_ev, rec_ev, ev_ms)
        return
    if args.workload == "newton":
        run_newton(args, rh, ctx, stream, new_ev, rec_ev, ev_ms)
        return
    if args.workload == "adjoint":
        run_adjoint(args, rh, ctx, stream, new_ev, rec_ev, ev_ms)
        return
    n = args.n
    dt = 1e-3
    t_setup = time.perf_counter()
    pb = rh.Problem(ctx, 3, (n, n, n), (1, 1, 1), 2, [0, 1], ["1", "0"])
    kappa = np.exp(0.5 * rh.philox_normal(20261201, 0, pb.n_elem))
    pb.set_conductivity(1.0, kappa=kappa)
    pb.set_mass(1.0, 1.0)
    t_setup = time.perf_counter() - t_setup
    nfree, nnz, slots = pb.n_free, pb.nnz, pb.slots
    col_mode, npat, _, _ = pb.layout()
    mode32, npat32, width32, slots32 = pb.layout_f32()
    layout32 = LAYOUTS[mode32]
    print(f"# setup {t_setup:.2f}s: nodes={pb.n_nodes} free={nfree} nnz={nnz} slots={slots} colours={pb.n_colours} "
          f"column mode={LAYOUTS[col_mode]} ({npat} patterns); fp32 copy: {layout32} ({npat32} patterns, "
          f"{slots32} slots)", file=sys.stderr, flush=True)

    results = {}
    # fp64 and mixed in the reference's element arithmetic (the FP64 loop is
    # bitwise the oracle's); then both once more with the uniform-geometry
    # arithmetic (the block mesh's element matrices from element 0's, within
    # a few ulps of the per-element ones), which removes the element pass;
    # and in the contracted arithmetic (per-point factors and fused
    # multiply-adds in every element, the fast assembly of a mesh of any shape)
    arith = {"fp64": rh.ARITH_REFERENCE, "mixed": rh.ARITH_REFERENCE,
             "fp64_uniform": rh.ARITH_UNIFORM, "mixed_uniform": rh.ARITH_UNIFORM,
             "fp64_contracted": rh.ARITH_CONTRACTED, "mixed_contracted": rh.ARITH_CONTRACTED}
    for mode in ("fp64", "mixed", "fp64_uniform", "mixed_uniform", "fp64_contracted", "mixed_contracted"):
        mixed = mode.startswith("mixed")
        pb.set_arithmetic(arith[mode])
        U = pb.vector(0.0)
        if args.warmup:
            pb.be_run(U, 0.0, args.warmup * dt, dt, tol=1e-8, max_iters=5000, mixed=mixed)
        U = pb.vector(0.0)
        ctx.sync()
        e0, e1 = new_ev(), new_ev()
        rec_ev(e0, stream)
        w0 = time.perf_counter()
        steps, iters = pb.be_run(U, 0.0, args.steps * dt, dt, tol=1e-8, max_iters=5000, mixed=mixed)
        rec_ev(e1, stream)
        total_ms = ev_ms(e0, e1)
        wall = time.perf_counter() - w0
        tm = dict(pb.last_timing_ms)
        its = int(np.sum(iters))
        results[mode] = dict(steps=steps, iters=its, iters_per_step=[int(v) for v in iters], total_ms=total_ms,
                             wall_s=wall, solve_ms=tm["solve"], assembly_ms=tm["assembly"],
                             dof_iters_per_s=nfree * its / (tm["solve"] * 1e-3),
                             dof_iters_per_s_step=nfree * its / (total_ms * 1e-3), U=U)
        print(f"# {mode}: steps={steps} iters={its} solve={tm['solve']:.1f}ms assembly={tm['assembly']:.1f}ms "
              f"total={total_ms:.1f}ms", file=sys.stderr, flush=True)
    pb.set_arithmetic(rh.ARITH_REFERENCE)
    Ud = results["fp64"]["U"].numpy()
    Um = results["mixed"]["U"].numpy()
    mixed_vs_fp64_err = float(np.max(np.abs(Um - Ud)) / max(np.max(np.abs(Ud)), 1e-300))
    # one step from T = 0 in each precision, for the comparison with the oracle
    one = {}
    if check is not None:
        for mode in ("fp64", "mixed"):
            U1 = pb.vector(0.0)
            _, it1 = pb.be_run(U1, 0.0, dt, dt, tol=1e-8, max_iters=5000, mixed=mode == "mixed")
            one[mode] = (U1.numpy(), int(it1[0]))
            U1.free()

    # SpMV roofline: the dominant kernel of each path, timed alone on the same matrix
    K = pb.matrix()
    pb.assemble_matrix(K, [(rh.DIFFUSION, 1.0), (rh.MASS | rh.SCALED, 1.0 / dt)], time=dt)
    K.to_f32()
    x64, y64 = pb.vector(1.0), pb.vector()
    x32, y32 = ctx.empty(nfree, np.float32), ctx.empty(nfree, np.float32)
    ctx.fill(x32, 1.0)
    # the mixed CG's own FP32 kernel reads x from a zero-padded vector
    lo, hi = pb.padding_f32()
    xpad = np.zeros(lo + nfree + hi, np.float32)
    xpad[lo:lo + nfree] = 1.0
    x32p = ctx.from_numpy(xpad)
    spmv = {}
    # algorithmic bytes of one product in the layout in use: values, the
    # column information (int32 columns, int16 offsets, or a 1-byte pattern
    # id per row), the row lengths where they are streamed, x once, y once;
    # the symmetric half storage holds (nnz + n) / 2 values, each read once
    # from HBM (the second reader of a pair finds it in L2 or not: the
    # measured traffic tells)
    def bytes_of(mode, vb):
        if mode == 4:  # no pattern id; the planes' zeros are overhead, not algorithmic
            return vb * ((nnz + nfree) // 2) + 2 * vb * nfree
        if mode == 3:
            return vb * ((nnz + nfree) // 2) + nfree + 2 * vb * nfree
        if mode == 2:
            return vb * nnz + nfree + 2 * vb * nfree
        if mode == 1:
            return (vb + 2) * nnz + 4 * (nfree + 1) + 2 * vb * nfree
        return (vb + 4) * nnz + 4 * (nfree + 1) + 2 * vb * nfree
    b64, b32 = bytes_of(col_mode, 8), bytes_of(mode32, 4)
    capture = hip_graph_capture()
    L = rh.lib()
    x32_at = C.c_void_p(x32p.ptr + 4 * lo)
    kernels = [("spmv_f64", lambda: rh.check(L.rsd_spmv_f64(ctx.h, K.h, rh.dptr(x64), rh.dptr(y64))), b64),
               ("spmv_f32", (lambda: rh.check(L.rsd_spmv_f32_padded(ctx.h, K.h, x32_at, rh.dptr(y32))))
                if mode32 == 4 else (lambda: rh.check(L.rsd_spmv_f32(ctx.h, K.h, rh.dptr(x32), rh.dptr(y32)))),
                b32)]
    st_points = pb.layout_stencil()[0]
    if st_points:
        # the FP64 CG's product: the stencil copy (values, a 4-byte position
        # mask per row, x once, y once; the copy's zeros are overhead)
        K.to_stencil()
        kernels.insert(1, ("spmv_f64_stencil",
                           lambda: rh.check(L.rsd_spmv_f64_stencil(ctx.h, K.h, rh.dptr(x64), rh.dptr(y64))),
                           8 * nnz + 4 * nfree + 16 * nfree))
    for name, fn, byts in kernels:
        for _ in range(3):
            fn()
        ctx.sync()
        replay = capture(fn, args.spmv_reps, stream)
        replay()
        ctx.sync()
        a, b = new_ev(), new_ev()
        rec_ev(a, stream)
        replay()
        rec_ev(b, stream)
        avg_ms = ev_ms(a, b) / args.spmv_reps
        spmv[name] = dict(avg_us=avg_ms * 1e3, bytes=byts, gbs=byts / (avg_ms * 1e-3) / 1e9,
                          frac=byts / (avg_ms * 1e-3) / 1e9 / PEAK_HBM_GBS)
        print(f"# {name}: {avg_ms*1e3:.1f} us/launch, {spmv[name]['gbs']:.0f} GB/s algorithmic", file=sys.stderr,
              flush=True)

    # matrix-free operator (FEMOperator): the same BE system solved with no
    # matrix, apply = element assembly on the operand, Jacobi from the
    # assembled diagonal; a bounded solve (one step's system)
    matrix_free = None
    if not args.no_matrix_free:
        be_forms = [(rh.DIFFUSION, 1.0), (rh.MASS | rh.SCALED, 1.0 / dt)]
        bx, yx = pb.vector(1.0), pb.vector()
        for _ in range(2):
            pb.fem_apply(yx, be_forms, bx)
        ctx.sync()
        a, b_ = new_ev(), new_ev()
        rec_ev(a, stream)
        reps = 5
        for _ in range(reps):
            pb.fem_apply(yx, be_forms, bx)
        rec_ev(b_, stream)
        apply_ms = ev_ms(a, b_) / reps
        # the factored product does every element's geometry per point, so
        # its time on this block mesh is its time on any Q1 mesh of the size
        other, cgs = {}, {}
        for name, ar in (("contracted", rh.ARITH_CONTRACTED), ("factored", rh.ARITH_FACTORED),
                         ("uniform", rh.ARITH_UNIFORM)):
            pb.set_arithmetic(ar)
            pb.fem_apply(yx, be_forms, bx)
            ctx.sync()
            a, b_ = new_ev(), new_ev()
            rec_ev(a, stream)
            for _ in range(reps):
                pb.fem_apply(yx, be_forms, bx)
            rec_ev(b_, stream)
            other[name] = ev_ms(a, b_) / reps
            if name != "contracted":
                xs = pb.vector(0.0)
                ctx.sync()
                a, b_ = new_ev(), new_ev()
                rec_ev(a, stream)
                r_ = pb.cg_fem(bx, xs, be_forms, tol=1e-8, max_iters=5000)
                rec_ev(b_, stream)
                cgs[name] = (int(r_.iterations), ev_ms(a, b_))
        pb.set_arithmetic(rh.ARITH_REFERENCE)
        xs = pb.vector(0.0)
        ctx.sync()
        a, b_ = new_ev(), new_ev()
        rec_ev(a, stream)
        rep = pb.cg_fem(bx, xs, be_forms, tol=1e-8, max_iters=5000)
        rec_ev(b_, stream)
        solve_ms = ev_ms(a, b_)
        matrix_free = {"apply_ms": apply_ms, "apply_over_spmv_f64": apply_ms * 1e3 / spmv["spmv_f64"]["avg_us"],
                       "apply_contracted_ms": other["contracted"], "apply_factored_ms": other["factored"],
                       "apply_factored_over_spmv_f64": other["factored"] * 1e3 / spmv["spmv_f64"]["avg_us"],
                       "apply_uniform_ms": other["uniform"],
                       "apply_uniform_over_spmv_f64": other["uniform"] * 1e3 / spmv["spmv_f64"]["avg_us"],
                       "factored_cg_iters": cgs["factored"][0], "factored_solve_ms": cgs["factored"][1],
                       "factored_dof_iters_per_s": nfree * cgs["factored"][0] / (cgs["factored"][1] * 1e-3),
                       "uniform_cg_iters": cgs["uniform"][0], "uniform_solve_ms": cgs["uniform"][1],
                       "uniform_dof_iters_per_s": nfree * cgs["uniform"][0] / (cgs["uniform"][1] * 1e-3),
                       "cg_iters": int(rep.iterations), "solve_ms": solve_ms,
                       "dof_iters_per_s": nfree * rep.iterations / (solve_ms * 1e-3)}
        print(f"# matrix-free: apply {apply_ms:.2f} ms, CG+Jacobi {rep.iterations} iterations in {solve_ms:.1f} ms; "
              f"contracted {other['contracted']:.3f} ms, factored {other['factored']:.3f} ms "
              f"(CG {cgs['factored'][1]:.1f} ms), uniform {other['uniform']:.3f} ms (CG {cgs['uniform'][1]:.1f} ms)",
              file=sys.stderr, flush=True)

    renumbered = None
    if not args.no_renumbered:
        renumbered = renumbered_workload(rh, ctx, stream, n, kappa, dt, capture, new_ev, rec_ev, ev_ms, args.spmv_reps,
                                         bytes_of)

    # calibration copy of known size (2^26 doubles = 512 MiB each way): the
    # FETCH_SIZE / WRITE_SIZE counters of this dispatch scale the SpMV's
    calib_n = 1 << 26
    ca, cb = ctx.empty(calib_n), ctx.empty(calib_n)
    ctx.fill(ca, 1.0)
    ctx.copy(ca, cb)
    ctx.sync()
    ca.free()
    cb.free()

    cpu = None
    if not args.no_cpu:
        b_host = np.ones(nfree)
        cb = cpu_baseline(pb, K, b_host, args.cpu_iters, kappa, dt)
        model, ncpu = cpu_info()
        ncores = len(_cores())
        if "ref_rate" in cb:
            cpu = {"value": cb["ref_rate"], "unit": "DOF-iters/s", "cores": 1, "kind": "reference",
                   "sample": f"{cb['ref_iters']} Jacobi-PCG iterations (differential: {args.cpu_iters} minus 20) of "
                             f"the reference's cg::Solver on the same {n}^3 BE operator ({nfree} DOFs), "
                             f"{cb['ref_s']:.1f} s pinned to core {cb['core']} of {ncores} ({model}, {ncpu} "
                             f"online); the reference's assembleMatrix of K + M/dt over {pb.n_elem} elements "
                             f"timed once",
                   "port_value": cb["port_rate"], "port_over_reference_time": cb["ref_rate"] / cb["port_rate"],
                   "assembly_elements_per_s": cb["ref_asm_elements_per_s"],
                   "ref_vs_gpu_matrix_max_rel_diff": cb["ref_vs_gpu_matrix_max_rel_diff"]}
        else:
            cpu = {"value": cb["port_rate"], "unit": "DOF-iters/s", "cores": 1, "kind": "port",
                   "sample": f"{cb['port_iters']} Jacobi-PCG iterations (differential) of the oracle's serial "
                             f"reference-order CG on the same {n}^3 BE operator ({nfree} DOFs), {cb['port_s']:.1f} s "
                             f"pinned to core {cb['core']} of {ncores} ({model})"}
        print(f"# cpu: {cpu['value']:.3e} DOF-iters/s ({cpu['kind']})", file=sys.stderr, flush=True)
    oracle_check = None
    if check is not None:
        oracle_check = finish_oracle_check(*check, one["fp64"][0], one["mixed"][0], one["fp64"][1], one["mixed"][1])
        print(f"# oracle check: {oracle_check}", file=sys.stderr, flush=True)

    mx, fp = results["mixed"], results["fp64"]
    fu, mu = results["fp64_uniform"], results["mixed_uniform"]
    fc, mc = results["fp64_contracted"], results["mixed_contracted"]
    steps = args.steps
    line = {
        "metric": "CG DOF-iters/s + SpMV HBM GB/s, mixed vs FP64 GPU at rel resid 1e-8",
        "value": mx["dof_iters_per_s"],
        "unit": "DOF-iters/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": mx["total_ms"] / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 inner CG / fp64 refinement",
        "data": "synthetic (seeded lognormal conductivity)",
        "config": {"workload": f"BE heat {n}^3 Q1 hex, dt=1e-3, CG+Jacobi to rel resid 1e-8", "dofs": nfree,
                   "nnz": nnz, "steps": steps},
        "fp64": {"dof_iters_per_s": fp["dof_iters_per_s"], "iters": fp["iters"], "solve_ms": fp["solve_ms"],
                 "assembly_ms": fp["assembly_ms"], "ms_per_step": fp["total_ms"] / steps},
        "mixed": {"dof_iters_per_s": mx["dof_iters_per_s"], "iters": mx["iters"], "solve_ms": mx["solve_ms"],
                  "assembly_ms": mx["assembly_ms"], "ms_per_step": mx["total_ms"] / steps},
        "uniform_geometry_arithmetic": {
            "fp64": {"iters": fu["iters"], "solve_ms": fu["solve_ms"], "assembly_ms": fu["assembly_ms"],
                     "ms_per_step": fu["total_ms"] / steps},
            "mixed": {"iters": mu["iters"], "solve_ms": mu["solve_ms"], "assembly_ms": mu["assembly_ms"],
                      "ms_per_step": mu["total_ms"] / steps, "dof_iters_per_s": mu["dof_iters_per_s"]},
            "step_time_fp64_over_mixed": fu["total_ms"] / mu["total_ms"],
            "elements_per_s": pb.n_elem * steps / (mu["assembly_ms"] * 1e-3),
            "fp64_vs_reference_fp64_max_rel_diff": float(np.max(np.abs(fu["U"].numpy() - Ud)) /
                                                         max(np.max(np.abs(Ud)), 1e-300))},
        "contracted_arithmetic": {
            "fp64": {"iters": fc["iters"], "solve_ms": fc["solve_ms"], "assembly_ms": fc["assembly_ms"],
                     "ms_per_step": fc["total_ms"] / steps},
            "mixed": {"iters": mc["iters"], "solve_ms": mc["solve_ms"], "assembly_ms": mc["assembly_ms"],
                      "ms_per_step": mc["total_ms"] / steps, "dof_iters_per_s": mc["dof_iters_per_s"]},
            "step_time_fp64_over_mixed": fc["total_ms"] / mc["total_ms"],
            "assembly_ms_per_step": mc["assembly_ms"] / steps,
            "elements_per_s": pb.n_elem * steps / (mc["assembly_ms"] * 1e-3),
            "fp64_vs_reference_fp64_max_rel_diff": float(np.max(np.abs(fc["U"].numpy() - Ud)) /
                                                         max(np.max(np.abs(Ud)), 1e-300))},
        "mixed_over_fp64": {"dof_iters_per_s": mx["dof_iters_per_s"] / fp["dof_iters_per_s"],
                            "solve_time": fp["solve_ms"] / mx["solve_ms"],
                            "step_time": fp["total_ms"] / mx["total_ms"]},
        # the per-step assembly is one pass over every element (K + M/dt with
        # M U_prev) plus the lift over the boundary elements
        "assembly": {"elements_per_s": pb.n_elem * steps / (mx["assembly_ms"] * 1e-3),
                     "ms_per_step": mx["assembly_ms"] / steps, "elements": pb.n_elem,
                     "work": "K + M/dt and M U_prev (one element pass), load, essential lift"},
        "mixed_vs_fp64_solution_max_rel_diff": mixed_vs_fp64_err,
        "oracle_check": oracle_check,
        "spmv": spmv,
        "matrix_free": matrix_free,
        "renumbered_mesh": renumbered,
        "roofline": {"bound": "hbm", "kernel": "k_spmv_f32<%d> (fp32 values, %s)" % (mode32, layout32),
                     "achieved": spmv["spmv_f32"]["gbs"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": spmv["spmv_f32"]["gbs"] / PEAK_HBM_GBS,
                     "traffic": measured_traffic(n, "k_spmv_f32<%d>" % mode32)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
