"""The reference's binary mesh files (PMSH v1, src/io/MeshIO.cpp:31-140),
read and written by the library's host code: no GPU is touched.

Fixtures: tests/golden/square_n{8,16}.pmsh are the reference's benchmark
meshes (benchmark/equation/heateq/fixtures/, written by its mesh
application from mesh_n{8,16}.yaml: block2d, Q1, [-1,1]^2)."""
import os

import numpy as np
import pytest

import oracle_py as O
import residuum_hip as rh

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("n", [8, 16])
def test_pmsh_fixture_is_the_block_mesh(n):
    """The reference's file holds BlockMesh2D(n, n, [-1,1]^2, Q1) exactly."""
    m = rh.read_pmsh(os.path.join(GOLDEN, "square_n%d.pmsh" % n))
    assert (m["pdim"], m["sdim"], m["family"], m["basis"]) == (2, 2, rh.PMSH_QUAD, rh.PMSH_LAGRANGE)
    assert (m["n_nodes"], m["n_elem"], m["npe"], m["faces"], m["orders"]) == ((n + 1) ** 2, n * n, 4, 4, [1, 1])
    xyz, ien, rng = O.block_mesh(2, (n, n), (-1.0, 1.0, -1.0, 1.0), (1, 1))
    assert np.array_equal(m["xyz"], xyz) and np.array_equal(m["ien"], ien) and np.array_equal(m["rng"], rng)


@pytest.mark.parametrize("n", [8, 16])
def test_pmsh_write_is_byte_exact(n, tmp_path):
    """Writing the block mesh reproduces the reference's file byte for byte."""
    xyz, ien, rng = O.block_mesh(2, (n, n), (-1.0, 1.0, -1.0, 1.0), (1, 1))
    out = tmp_path / "m.pmsh"
    rh.write_pmsh(str(out), dict(pdim=2, sdim=2, family=rh.PMSH_QUAD, n_nodes=(n + 1) ** 2, n_elem=n * n, npe=4,
                                 faces=4, orders=[1, 1], xyz=xyz, ien=ien, rng=rng))
    with open(os.path.join(GOLDEN, "square_n%d.pmsh" % n), "rb") as f:
        assert out.read_bytes() == f.read()


@pytest.mark.parametrize("dim,n,p", [(3, (3, 2, 4), (1, 1, 1)), (2, (3, 5), (2, 2)), (3, (2, 2, 1), (2, 2, 2))])
def test_pmsh_roundtrip(dim, n, p, tmp_path):
    ext = (0.0, 1.0) * dim
    xyz, ien, rng = O.block_mesh(dim, n, ext, p)
    npe = int(np.prod([q + 1 for q in p]))
    mesh = dict(pdim=dim, sdim=dim, family=rh.PMSH_QUAD if dim == 2 else rh.PMSH_HEX, n_nodes=xyz.size // dim,
                n_elem=ien.size // npe, npe=npe, faces=2 * dim, orders=list(p), xyz=xyz, ien=ien, rng=rng)
    out = str(tmp_path / "m.pmsh")
    rh.write_pmsh(out, mesh)
    back = rh.read_pmsh(out)
    for k in ("pdim", "sdim", "family", "n_nodes", "n_elem", "npe", "faces", "orders"):
        assert back[k] == mesh[k]
    assert np.array_equal(back["xyz"], xyz) and np.array_equal(back["ien"], ien) and np.array_equal(back["rng"], rng)


def test_pmsh_errors(tmp_path):
    bad = tmp_path / "bad.pmsh"
    bad.write_bytes(b"NOTAPMSHFILE" * 4)
    with pytest.raises(rh.RsdError, match="magic"):
        rh.read_pmsh(str(bad))
    with open(os.path.join(GOLDEN, "square_n8.pmsh"), "rb") as f:
        data = f.read()
    cut = tmp_path / "cut.pmsh"
    cut.write_bytes(data[:len(data) // 2])
    with pytest.raises(rh.RsdError, match="end of file"):
        rh.read_pmsh(str(cut))
    with pytest.raises(rh.RsdError, match="cannot open"):
        rh.read_pmsh(str(tmp_path / "missing.pmsh"))
    xyz, ien, rng = O.block_mesh(2, (2, 2), (0.0, 1.0, 0.0, 1.0), (1, 1))
    with pytest.raises(rh.RsdError, match="nodes per element"):
        rh.write_pmsh(str(tmp_path / "w.pmsh"), dict(pdim=2, sdim=2, family=0, n_nodes=9, n_elem=4, npe=4, faces=4,
                                                     orders=[2, 1], xyz=xyz, ien=ien, rng=rng))


# ---------------------------------------------------------------------------
# the command line front end (configuration files of the reference's schema)
# ---------------------------------------------------------------------------
def test_mesh_app_writes_the_reference_file(tmp_path):
    """`mesh mesh_n8.yaml` (the reference's benchmark mesh configuration)
    writes square_n8.pmsh byte for byte as the reference's mesh app did."""
    from residuum_hip import cli
    cfg = tmp_path / "mesh.yaml"
    cfg.write_text(open(os.path.join(GOLDEN, "mesh_n8.yaml")).read())
    out = cli.run_mesh(str(cfg), quiet=True)
    with open(os.path.join(GOLDEN, "square_n8.pmsh"), "rb") as f:
        assert open(out, "rb").read() == f.read()


def test_heat_config_parse():
    """The reference's benchmark configuration reads into the settings its
    parsers give (LinearSolverConfigParser.cpp, DiscretizationConfigParser.cpp, ...)."""
    from residuum_hip import cli
    cfg = cli.parse_heat_config(os.path.join(GOLDEN, "heateq_square_n16.yaml"))
    assert cfg["mesh"].endswith("square_n16.pmsh") and cfg["quadrature"] == (2, 2, 2)
    assert cfg["conductivity"]["type"] == "constant" and cfg["conductivity"]["value"] == 1.0
    assert cfg["bc_tags"] == [0, 1, 2, 3] and cfg["bc_exprs"] == ["(1-x)*(1-y)"] * 4
    assert cfg["source"] == ("expression", "0.0") and cfg["initial"] == ("expression", "0.0")
    assert cfg["driver"] == "steady"
    assert cfg["linear"] == dict(type="cg", tolerance=1e-12, max_iterations=10000, krylov_dim=50, operator="fem",
                                 preconditioner="identity", precision="fp64", orthogonalization="mgs")


def test_heat_config_arithmetic(tmp_path):
    """`discretization: arithmetic:` selects the element arithmetic; the
    reference's order is the default."""
    import yaml
    from residuum_hip import cli
    d = yaml.safe_load(open(os.path.join(GOLDEN, "heateq_square_n16.yaml")))
    p = tmp_path / "c.yaml"
    p.write_text(yaml.safe_dump(d))
    assert cli.parse_heat_config(str(p))["arithmetic"] == "reference"
    d["discretization"]["arithmetic"] = "uniform"
    p.write_text(yaml.safe_dump(d))
    assert cli.parse_heat_config(str(p))["arithmetic"] == "uniform"


def test_heat_config_precision(tmp_path):
    """`precision: mixed` on a linear-solver node selects the mixed solves;
    fp64 is the default."""
    import yaml
    from residuum_hip import cli
    d = yaml.safe_load(open(os.path.join(GOLDEN, "heateq_square_n16.yaml")))
    d["solver"]["linear"]["operator"]["type"] = "csr"
    d["solver"]["linear"]["precision"] = "mixed"
    p = tmp_path / "c.yaml"
    p.write_text(yaml.safe_dump(d))
    assert cli.parse_heat_config(str(p))["linear"]["precision"] == "mixed"


def test_heat_config_orthogonalization(tmp_path):
    """`orthogonalization: cgs2` on a gmres node selects classical
    Gram-Schmidt with reorthogonalisation; the reference's mgs is the default."""
    import yaml
    from residuum_hip import cli
    d = yaml.safe_load(open(os.path.join(GOLDEN, "heateq_square_n16.yaml")))
    d["solver"]["linear"].update(type="gmres", orthogonalization="cgs2")
    d["solver"]["linear"]["operator"]["type"] = "csr"
    p = tmp_path / "c.yaml"
    p.write_text(yaml.safe_dump(d))
    assert cli.parse_heat_config(str(p))["linear"]["orthogonalization"] == "cgs2"


@pytest.mark.parametrize("edit,msg", [
    (lambda d: d.update(backend="cpu"), "backend"),
    (lambda d: d["solver"]["linear"].update(type="bicgstab"), "bicgstab"),
    (lambda d: d["boundary_conditions"][0].update(type="flux", forms=["flux_bc"]), "one expression per spatial"),
    (lambda d: d["boundary_conditions"][0].update(type="flux", forms=["flux_bc"], read={"mode": "file", "file": "q"}),
     "nodal file"),
    (lambda d: d["boundary_conditions"][0].update(type="flux"), "other than expressions"),
    (lambda d: d["solver"]["linear"]["preconditioner"].update(type="ilu"), "preconditioner"),
    (lambda d: d["solver"]["driver"].update(type="pseudo_transient"), "driver"),
    (lambda d: d.pop("solver"), "solver"),
    (lambda d: d["solver"]["linear"].update(precision="fp16"), "precision"),
    (lambda d: d["solver"]["linear"].update(precision="mixed"), "csr operator"),
    (lambda d: d["discretization"].update(arithmetic="fast"), "arithmetic"),
    (lambda d: d["solver"]["linear"].update(orthogonalization="householder"), "orthogonalization"),
    (lambda d: d["solver"]["linear"].update(orthogonalization="cgs2"), "applies to gmres"),
])
def test_heat_config_refusals(edit, msg, tmp_path):
    """What the device path does not run is refused with a message."""
    import yaml
    from residuum_hip import cli
    d = yaml.safe_load(open(os.path.join(GOLDEN, "heateq_square_n16.yaml")))
    edit(d)
    p = tmp_path / "c.yaml"
    p.write_text(yaml.safe_dump(d))
    with pytest.raises(cli.ConfigError, match=msg):
        cli.parse_heat_config(str(p))


MONITORS = [  # the monitor blocks of the reference's examples/heateq configurations
    {"name": "net_influx", "quantity": "heat_flux", "flux_method": "weak", "reduction": "integral",
     "terms": [{"boundary": b, "coefficient": -1.0} for b in range(4)],
     "output": {"console": True, "file": "output/monitors/net_influx.csv"}},
    {"name": "wall_flux", "quantity": "heat_flux", "flux_method": "strong", "reduction": "integral",
     "terms": [{"boundary": 3}], "output": {"console": True, "file": "output/monitors/wall_flux.csv"}},
    {"name": "inlet_flux_avg", "quantity": "heat_flux", "reduction": "average", "terms": [{"boundary": 1}]},
]


def test_heat_config_monitors(tmp_path):
    """MonitorConfigParser.cpp: flux_method defaults to weak, coefficient to
    1, console to true, file to none."""
    import yaml
    from residuum_hip import cli
    d = yaml.safe_load(open(os.path.join(GOLDEN, "heateq_square_n16.yaml")))
    d["monitors"] = MONITORS
    p = tmp_path / "c.yaml"
    p.write_text(yaml.safe_dump(d))
    mons = cli.parse_heat_config(str(p))["monitors"]
    assert [m["name"] for m in mons] == ["net_influx", "wall_flux", "inlet_flux_avg"]
    assert mons[0]["terms"] == [(b, -1.0) for b in range(4)] and mons[0]["method"] == "weak"
    assert mons[0]["file"] == str(tmp_path / "output/monitors/net_influx.csv")
    assert mons[1]["method"] == "strong" and mons[1]["terms"] == [(3, 1.0)]
    assert mons[2] == dict(name="inlet_flux_avg", method="weak", reduction="average", terms=[(1, 1.0)],
                           console=True, file="")


@pytest.mark.parametrize("edit,msg", [
    (lambda m: m.update(flux_method="pointwise"), "unknown method"),
    (lambda m: m.update(reduction="max"), "unknown reduction"),
    (lambda m: m.pop("reduction"), "reduction"),
    (lambda m: m.update(quantity="temperature"), "Unknown monitor quantity"),
    (lambda m: m.pop("terms"), "missing required 'terms'"),
    (lambda m: m["terms"][0].pop("boundary"), "boundary"),
])
def test_heat_config_monitor_errors(edit, msg, tmp_path):
    import copy
    import yaml
    from residuum_hip import cli
    d = yaml.safe_load(open(os.path.join(GOLDEN, "heateq_square_n16.yaml")))
    m = copy.deepcopy(MONITORS[0])
    edit(m)
    d["monitors"] = [m]
    p = tmp_path / "c.yaml"
    p.write_text(yaml.safe_dump(d))
    with pytest.raises(cli.ConfigError, match=msg):
        cli.parse_heat_config(str(p))
