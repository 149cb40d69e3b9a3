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
            {"boundary": 0, "type": "value", "read": {"mode": "expression", "expression": "0.0"}},
            {"boundary": 1, "type": "value", "read": {"mode": "expression", "expression": "1.0"}}],
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
