"""The command line front end on the device: the reference's configuration
files run end to end (python -m residuum_hip heateq config.yaml)."""
import os
import shutil

import numpy as np
import pytest
import yaml

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)


def nodes_xy(ctx, mesh_path, nq=2, tags=(0, 1, 2, 3), exprs=("0",) * 4):
    import residuum_hip as rh
    pb = rh.Problem.from_pmsh(ctx, mesh_path, nq, list(tags), list(exprs))
    xyz, _, _ = pb.mesh()
    _, a2t, _ = pb.dofs()
    return xyz.reshape(-1, pb.dim)[a2t]


def write_cfg(tmp_path, d, name="config.yaml"):
    p = tmp_path / name
    p.write_text(yaml.safe_dump(d))
    return str(p)


def bench_config(tmp_path):
    shutil.copy(os.path.join(GOLDEN, "square_n16.pmsh"), tmp_path / "square_n16.pmsh")
    return yaml.safe_load(open(os.path.join(GOLDEN, "heateq_square_n16.yaml")))


@pytest.mark.parametrize("operator,precond", [("fem", "identity"), ("csr", "identity"), ("csr", "jacobi")])
def test_steady_bilinear(ctx, tmp_path, operator, precond):
    """The reference's benchmark configuration (matrix-free CG, 1e-12) and
    its CSR / Jacobi variants: the bilinear solution to 1e-10."""
    from residuum_hip import cli
    d = bench_config(tmp_path)
    d["solver"]["linear"]["operator"]["type"] = operator
    d["solver"]["linear"]["preconditioner"]["type"] = precond
    say("steady", operator, precond)
    s = cli.run_heat(write_cfg(tmp_path, d), quiet=True, ctx=ctx)
    assert s["converged"] and abs(s["final_residual"] - 1e-12) <= 1e-11
    xy = nodes_xy(ctx, str(tmp_path / "square_n16.pmsh"))
    assert np.max(np.abs(s["U"] - (1 - xy[:, 0]) * (1 - xy[:, 1]))) <= 1e-10


@pytest.mark.parametrize("solver", ["cg", "gmres"])
def test_steady_mixed_precision(ctx, tmp_path, solver):
    """`precision: mixed` runs the steady solve with FP32 inner solves and
    FP64 refinement: the bilinear solution to 1e-10, as in FP64."""
    from residuum_hip import cli
    d = bench_config(tmp_path)
    d["solver"]["linear"].update(type=solver, precision="mixed", krylov_dim=30)
    d["solver"]["linear"]["operator"]["type"] = "csr"
    d["solver"]["linear"]["preconditioner"]["type"] = "jacobi"
    s = cli.run_heat(write_cfg(tmp_path, d), quiet=True, ctx=ctx)
    say("steady mixed", solver, s["iterations"], s["final_residual"])
    assert s["converged"]
    xy = nodes_xy(ctx, str(tmp_path / "square_n16.pmsh"))
    assert np.max(np.abs(s["U"] - (1 - xy[:, 0]) * (1 - xy[:, 1]))) <= 1e-10


@pytest.mark.parametrize("ortho", ["mgs", "cgs2"])
def test_steady_gmres_orthogonalization(ctx, tmp_path, ortho):
    """A steady GMRES(30) + Jacobi solve with either orthogonalisation
    (`orthogonalization: cgs2` is classical Gram-Schmidt reorthogonalised
    once): the bilinear solution to 1e-10."""
    import residuum_hip as rh
    from residuum_hip import cli
    d = bench_config(tmp_path)
    d["solver"]["linear"].update(type="gmres", krylov_dim=30, orthogonalization=ortho)
    d["solver"]["linear"]["operator"]["type"] = "csr"
    d["solver"]["linear"]["preconditioner"]["type"] = "jacobi"
    s = cli.run_heat(write_cfg(tmp_path, d), quiet=True, ctx=ctx)
    say("steady gmres", ortho, s["iterations"], s["final_residual"])
    assert s["converged"]
    xy = nodes_xy(ctx, str(tmp_path / "square_n16.pmsh"))
    assert np.max(np.abs(s["U"] - (1 - xy[:, 0]) * (1 - xy[:, 1]))) <= 1e-10
    with pytest.raises(rh.RsdError, match="orthogonalisation"):
        ctx.set_gmres_orthogonalization(7)


def test_transient_mixed_precision(ctx, tmp_path):
    """The Backward Euler loop on the device with `precision: mixed` agrees
    with the FP64 loop to the solver tolerance."""
    from residuum_hip import cli
    d = bench_config(tmp_path)
    for bc in d["boundary_conditions"]:
        bc["read"]["expression"] = "100.0"
    d["physics"]["models"]["source"]["read"]["expression"] = "10*x*y"
    d["solver"]["driver"]["type"] = "transient"
    d["solver"]["timestepper"] = {"type": "backward_euler", "t0": 0.0, "tf": 0.1, "step_size": {"mode": "constant", "dt": 0.02}}
    d["solver"]["linear"] = {"type": "cg", "operator": {"type": "csr"}, "preconditioner": {"type": "jacobi"},
                             "tolerance": 1e-10, "max_iterations": 5000}
    s64 = cli.run_heat(write_cfg(tmp_path, d), quiet=True, ctx=ctx)
    d["solver"]["linear"]["precision"] = "mixed"
    smx = cli.run_heat(write_cfg(tmp_path, d, "mixed.yaml"), quiet=True, ctx=ctx)
    assert s64["steps"] == smx["steps"] == 5
    assert np.max(np.abs(smx["U"] - s64["U"])) <= 1e-8 * np.max(np.abs(s64["U"]))


def test_transient_paths_agree(ctx, tmp_path):
    """Backward Euler from a configuration: the whole loop on the device
    (CSR + CG) equals Problem.be_run bit for bit; the matrix-free operator
    run one solve per step agrees to the solver tolerance; the output step
    files hold the nodal fields."""
    import residuum_hip as rh
    from residuum_hip import cli
    d = bench_config(tmp_path)
    for bc in d["boundary_conditions"]:
        bc["read"]["expression"] = "100.0"
    d["materials"]["density"] = {"value": 2.0, "unit": "kg/m^3"}
    d["materials"]["specific_heat"] = {"value": 1.5, "unit": "J/(kg*K)"}
    d["physics"]["models"]["source"]["read"]["expression"] = "10*x*y"
    d["solver"]["driver"]["type"] = "transient"
    d["solver"]["timestepper"] = {"type": "backward_euler", "t0": 0.0, "tf": 0.1, "step_size": {"mode": "constant", "dt": 0.02}}
    d["solver"]["linear"] = {"type": "cg", "operator": {"type": "csr"}, "preconditioner": {"type": "jacobi"},
                             "tolerance": 1e-12, "max_iterations": 5000}
    s_dev = cli.run_heat(write_cfg(tmp_path, d), quiet=True, ctx=ctx)
    assert s_dev["steps"] == 5
    pb = rh.Problem.from_pmsh(ctx, str(tmp_path / "square_n16.pmsh"), 2, [0, 1, 2, 3], ["100.0"] * 4)
    pb.set_mass(2.0, 1.5)
    U = pb.vector(0.0)
    steps, _ = pb.be_run(U, 0.0, 0.1, 0.02, src="10*x*y", tol=1e-12, max_iters=5000)
    assert steps == 5 and np.array_equal(U.numpy(), s_dev["U"])
    d["solver"]["linear"]["operator"]["type"] = "fem"
    d["output"] = {"directory": "out", "format": "vtu", "write_frequency": 1, "prefix": "T"}
    s_fem = cli.run_heat(write_cfg(tmp_path, d, "fem.yaml"), quiet=True, ctx=ctx)
    say("fem iterations", s_fem["iterations"])
    assert s_fem["steps"] == 5
    assert np.max(np.abs(s_fem["U"] - s_dev["U"])) <= 1e-9 * np.max(np.abs(s_dev["U"]))
    nodal, t = pb.read_pndf(str(tmp_path / "out" / "T_000005.pndf"))
    assert abs(t - 0.1) < 1e-12
    assert np.array_equal(nodal, pb.nodal_field(ctx.from_numpy(s_fem["U"]), 0.1))
    # the element arithmetic of the discretization node: the uniform-geometry
    # pair on this block mesh, through both operators, within the solver
    # tolerance of the reference order
    d.pop("output")
    d["discretization"]["arithmetic"] = "uniform"
    for op, ref in (("csr", s_dev["U"]), ("fem", s_fem["U"])):
        d["solver"]["linear"]["operator"]["type"] = op
        s_u = cli.run_heat(write_cfg(tmp_path, d, "uniform_%s.yaml" % op), quiet=True, ctx=ctx)
        say(op, "uniform arithmetic: max rel diff", np.max(np.abs(s_u["U"] - ref)) / np.max(np.abs(ref)))
        assert s_u["steps"] == 5 and np.max(np.abs(s_u["U"] - ref)) <= 1e-11 * np.max(np.abs(ref))
    assert os.path.getsize(tmp_path / "out" / "T_000005.vtu") > 1000


def test_steady_nonlinear_kirchhoff(ctx, tmp_path):
    """A temperature-dependent conductivity runs Newton
    (HeatEquationSteadyNonlinearConductivity2D.cpp:29-64: k(T) = 1 + 0.5 T,
    T = 0 at x = 0, 1 at x = 1): the Kirchhoff closed form to 1e-6."""
    from residuum_hip import cli
    (tmp_path / "mesh.yaml").write_text(yaml.safe_dump({
        "mesh": {"type": "block2d", "px": 1, "py": 1, "nx": 40, "ny": 4, "xmin": 0.0, "xmax": 1.0, "ymin": 0.0,
                 "ymax": 1.0}, "output": {"output_file": "strip.pmsh"}}))
    cli.run_mesh(str(tmp_path / "mesh.yaml"), quiet=True)
    d = {
        "mesh": {"file": "strip.pmsh"},
        "discretization": {"quadrature": {"xi": 2, "eta": 2}},
        "materials": {"conductivity": {"type": "temperature_dependent_isotropic", "value_expression": "1.0 + 0.5*T",
                                       "gradient_expression": "0.5", "unit": "W/(m*K)"}},
        "boundary_conditions": [
            {"boundary": 0, "type": "value", "read": {"mode": "expression", "expression": "0.0", "unit": "K"},
             "forms": ["value_bc"], "model": "constant"},
            {"boundary": 1, "type": "value", "read": {"mode": "expression", "expression": "1.0", "unit": "K"},
             "forms": ["value_bc"], "model": "constant"}],
        "solver": {"driver": {"type": "steady"},
                   "nonlinear": {"type": "newton", "absolute_tolerance": 1e-10, "relative_tolerance": 1e-10,
                                 "max_iterations": 30,
                                 "linear_solver": {"type": "cg", "preconditioner": {"type": "jacobi"},
                                                   "tolerance": 1e-12, "max_iterations": 5000}}},
    }
    s = cli.run_heat(write_cfg(tmp_path, d), quiet=True, ctx=ctx)
    assert s["converged"] and s["newton_iterations"] > 0
    x = nodes_xy(ctx, str(tmp_path / "strip.pmsh"), tags=(0, 1), exprs=("0", "0"))[:, 0]
    theta = (-1.0 + np.sqrt(1.0 + 2 * 0.5 * (1.0 + 0.25) * x)) / 0.5
    assert np.max(np.abs(s["U"] - theta)) < 1e-6


def test_monitors_steady_and_transient(ctx, tmp_path):
    """Monitors from a configuration (HeatProblem.tpp:133-167, 446-460): the
    strong flux of the bilinear solution (1-x)(1-y) on [-1,1]^2 is -2, 2,
    -2, 2 on the four sides (average: the integral over the length 2); the
    weak values equal the device reaction of the final state, combined
    with the coefficients; CSV rows (tick, time, value).  A transient run
    reports the initial condition and every step."""
    import residuum_hip as rh
    from residuum_hip import cli
    d = bench_config(tmp_path)
    d["solver"]["linear"]["operator"]["type"] = "csr"
    d["monitors"] = [
        {"name": "left_strong", "quantity": "heat_flux", "flux_method": "strong", "reduction": "integral",
         "terms": [{"boundary": 0}], "output": {"console": False, "file": "mon/left_strong.csv"}},
        {"name": "net_strong", "quantity": "heat_flux", "flux_method": "strong", "reduction": "integral",
         "terms": [{"boundary": b} for b in range(4)]},
        {"name": "top_avg", "quantity": "heat_flux", "flux_method": "strong", "reduction": "average",
         "terms": [{"boundary": 3}]},
        {"name": "net_influx", "quantity": "heat_flux", "reduction": "integral",
         "terms": [{"boundary": b, "coefficient": -1.0} for b in range(4)],
         "output": {"file": "mon/net_influx.csv"}},
        {"name": "right_weak_avg", "quantity": "heat_flux", "flux_method": "weak", "reduction": "average",
         "terms": [{"boundary": 1, "coefficient": 2.0}]},
    ]
    s = cli.run_heat(write_cfg(tmp_path, d), quiet=True, ctx=ctx)
    mv = {k: v[0][2] for k, v in s["monitors"].items()}
    say("monitors", mv)
    assert abs(mv["left_strong"] + 2.0) <= 1e-10 and abs(mv["net_strong"]) <= 1e-10
    assert abs(mv["top_avg"] - 1.0) <= 1e-10
    pb = rh.Problem.from_pmsh(ctx, str(tmp_path / "square_n16.pmsh"), 2, [0, 1, 2, 3], ["(1-x)*(1-y)"] * 4)
    dU = ctx.from_numpy(s["U"])
    res, meas = pb.boundary_reaction([0, 1, 2, 3], dU, src="0.0")
    want = 0.0
    for b in range(4):
        want += -1.0 * res[b]
    assert mv["net_influx"] == want
    assert mv["right_weak_avg"] == 2.0 * (res[1] / meas[1])
    rows = open(tmp_path / "mon" / "left_strong.csv").read().splitlines()
    assert rows[0] == "tick,time,left_strong [W]" and rows[1].startswith("0,0,")
    assert abs(float(rows[1].split(",")[2]) + 2.0) < 1e-5
    # transient: ticks 0..5 at t0 + k dt, the weak group with the step's rate
    d["solver"]["driver"]["type"] = "transient"
    d["solver"]["timestepper"] = {"type": "backward_euler", "t0": 0.0, "tf": 0.1,
                                  "step_size": {"mode": "constant", "dt": 0.02}}
    d["solver"]["linear"] = {"type": "cg", "operator": {"type": "csr"}, "preconditioner": {"type": "jacobi"},
                             "tolerance": 1e-12, "max_iterations": 5000}
    st = cli.run_heat(write_cfg(tmp_path, d, "tr.yaml"), quiet=True, ctx=ctx)
    ticks = st["monitors"]["net_influx"]
    assert [t for t, _, _ in ticks] == list(range(6))
    assert np.allclose([t for _, t, _ in ticks], [0.02 * k for k in range(6)], atol=1e-15, rtol=0)
    res_t, _ = pb.boundary_reaction([0, 1, 2, 3], ctx.from_numpy(st["U"]), Udot=pb.vector(0.0), src="0.0", time=0.1)
    want = 0.0
    for b in range(4):
        want += -1.0 * res_t[b]
    assert ticks[-1][2] == want
    rows = open(tmp_path / "mon" / "net_influx.csv").read().splitlines()
    assert len(rows) == 7 and rows[0] == "tick,time,net_influx [W]"


@pytest.mark.parametrize("operator", ["csr", "fem"])
def test_flux_boundary_condition(ctx, tmp_path, operator):
    """A flux_bc (BoundaryConditionConfigParser.cpp:43-50, applied by
    applyNaturalBCs) on x = 1 in place of the value there: with the flux
    K grad T of T = (1-x)(1-y) the discrete solution is still that bilinear
    field (the face integral of the flux is exact with 2 Gauss points), both
    steady and as a steady state kept by the Backward Euler loop."""
    from residuum_hip import cli
    d = bench_config(tmp_path)
    d["solver"]["linear"]["operator"]["type"] = operator
    d["solver"]["linear"]["preconditioner"]["type"] = "jacobi"
    for bc in d["boundary_conditions"]:
        if bc["boundary"] == 1:
            bc.update(type="flux", forms=["flux_bc"])
            bc["read"] = {"mode": "expression", "expression": ["-(1-y)", "-(1-x)"], "unit": "W/m^2"}
    s = cli.run_heat(write_cfg(tmp_path, d), quiet=True, ctx=ctx)
    xy = nodes_xy(ctx, str(tmp_path / "square_n16.pmsh"), tags=(0, 2, 3), exprs=("0",) * 3)
    exact = (1 - xy[:, 0]) * (1 - xy[:, 1])
    say("flux steady", operator, s["iterations"], np.max(np.abs(s["U"] - exact)))
    assert s["converged"] and s["U"].size == exact.size
    assert np.max(np.abs(s["U"] - exact)) <= 1e-9
    d["solver"]["driver"]["type"] = "transient"
    d["solver"]["timestepper"] = {"type": "backward_euler", "t0": 0.0, "tf": 0.06,
                                  "step_size": {"mode": "constant", "dt": 0.02}}
    d["initial_condition"] = {"read": {"mode": "expression", "expression": "(1-x)*(1-y)", "unit": "K"}}
    d["materials"]["density"] = {"value": 1.0, "unit": "kg/m^3"}
    d["materials"]["specific_heat"] = {"type": "constant", "value": 1.0, "unit": "J/(kg*K)"}
    s = cli.run_heat(write_cfg(tmp_path, d, "transient.yaml"), quiet=True, ctx=ctx)
    say("flux transient", operator, s["iterations"], np.max(np.abs(s["U"] - exact)))
    assert s["converged"] and s["steps"] == 3
    assert np.max(np.abs(s["U"] - exact)) <= 1e-9
