"""Command line front end: the reference's `residuum heateq config.yaml` and
`residuum mesh mesh.yaml` (src/application/ResiduumApplication.cpp:9-18)
with the work on the device.

    python -m residuum_hip heateq config.yaml
    python -m residuum_hip mesh mesh.yaml

The configuration files are the reference's own YAML schema; the parts read
here follow its parsers key for key (src/solver/parser/*,
src/application/heateq/parser/*).  `backend:` is optional; when present it
must be `hip` (the reference's parser knows cpu and cuda,
BackendConfigParser.cpp:6-14).  What the device path does not run is refused
with a message, never run on the host: fluxes from nodal files or with the
Newton solver, boundary values or sources other than expressions or
(sources) nodal files,
weak flux monitors together with a nodal-file source, non-Lagrange meshes,
the pseudo-transient driver, and time steppers other than Backward Euler
with a constant step.

Boundary heat-flux monitors (MonitorConfigParser.cpp, HeatProblem.tpp:
133-167, 446-460) run on the device: monitors with the same (method,
reduction) share one evaluation over the union of their tags; each reports
sum(coefficient x result(tag)), the result divided by the tags' face measure
for the average reduction; console lines and CSV rows (tick, time, value)
after the steady solve (tick 0, time 0) or at the initial condition and
after every step of a transient run.

The flow is HeatDispatcher::run (src/application/heateq/HeatDispatcher.cpp):
read the mesh, build the problem, then the steady stage (linear:
SteadyStage.hpp:60-84; nonlinear conductivity: Newton, Newton.hpp:32-83) or
the transient Backward Euler loop (BackwardEulerStage.hpp:71-163) with the
CSR operator and CG (rsd_be_run, the whole loop on the device) or, for the
FEM operator or GMRES, one device solve per step.

A linear-solver node may carry `precision: mixed` (a device extension; the
default `fp64` is the reference's behaviour): FP32 inner solves with FP64
iterative refinement -- rsd_cg_solve_mixed / rsd_gmres_solve_mixed for a
steady solve, the mixed Backward Euler loop, and Newton's mixed inner
solves (GMRES, or CG on a full FP32 copy of the Jacobian).  It needs the
csr operator.

A gmres node may carry `orthogonalization: cgs2` (a device extension; the
default `mgs` is the reference's modified Gram-Schmidt, bitwise): classical
Gram-Schmidt reorthogonalised once, three folds per Arnoldi step instead of
one per basis vector, for the GMRES solves (steady and Newton's inner
solves, FP64 or the mixed solver's FP32 inner GMRES).

`discretization: arithmetic:` (a device extension; default `reference`, the
reference's operation order, bitwise) selects the element arithmetic of the
Backward Euler form pair: `contracted` (fused multiply-adds), `uniform`
(on a block mesh, every element matrix from element 0's: no element pass)
or `factored` (on Q1 meshes of any shape, the fem operator's product
without element matrices); all within a few 1e-15 of the reference order.
"""
import argparse
import contextlib
import os
import sys
import time as _time

import numpy as np

from . import (ARITH_CONTRACTED, ARITH_FACTORED, ARITH_REFERENCE, ARITH_UNIFORM, DIFFUSION, MASS, SCALED, PC_IDENTITY, PC_JACOBI,
               ORTHO_CGS2, ORTHO_MGS,
               SEL_VALUE, PMSH_HEX, PMSH_LAGRANGE, PMSH_QUAD, Ctx, Problem, RsdError, block_mesh, read_pmsh,
               write_pmsh)

ORTHO = {"mgs": ORTHO_MGS, "cgs2": ORTHO_CGS2}
ARITHMETIC = {"reference": ARITH_REFERENCE, "contracted": ARITH_CONTRACTED, "uniform": ARITH_UNIFORM,
              "factored": ARITH_FACTORED}


class ConfigError(RuntimeError):
    pass


def _req(node, key, where):
    if not isinstance(node, dict) or key not in node:
        raise ConfigError(f"{where}: missing required '{key}'")
    return node[key]


def _opt(node, key, default):
    return node.get(key, default) if isinstance(node, dict) else default


# ---------------------------------------------------------------------------
# configuration (the reference's parsers, the subset the device runs)
# ---------------------------------------------------------------------------
def _read_field(node, where, allow_file=False):
    """NodalFieldReadConfigParser: {mode: expression, expression} or
    {mode: file, file}."""
    read = _req(node, "read", where)
    mode = _req(read, "mode", where + ".read")
    if mode == "expression":
        return ("expression", str(_req(read, "expression", where + ".read")))
    if mode == "file" and allow_file:
        return ("file", str(_req(read, "file", where + ".read")))
    raise ConfigError(f"{where}: read mode '{mode}' is not supported on the device"
                      + (" (expression or file)" if allow_file else " (expression)"))


def _linear(node, where):
    """LinearSolverConfigParser.cpp:44-85."""
    typ = _req(node, "type", where)
    if typ not in ("cg", "gmres"):
        raise ConfigError(f"{where}: linear solver '{typ}' is not available on the device (cg, gmres)")
    op = _opt(node, "operator", None)
    pc = _opt(node, "preconditioner", None)
    cfg = dict(type=typ, tolerance=float(_opt(node, "tolerance", 1e-10)),
               max_iterations=int(_opt(node, "max_iterations", 1000)),
               krylov_dim=int(_opt(node, "krylov_dim", 50)),
               operator=_req(op, "type", where + ".operator") if op else "csr",
               preconditioner=_req(pc, "type", where + ".preconditioner") if pc else "identity",
               precision=str(_opt(node, "precision", "fp64")),
               orthogonalization=str(_opt(node, "orthogonalization", "mgs")))
    if cfg["operator"] not in ("csr", "fem"):
        raise ConfigError(f"Unknown linear solver operator type: {cfg['operator']}. Valid options: csr, fem")
    if cfg["precision"] not in ("fp64", "mixed"):
        raise ConfigError(f"{where}: unknown precision '{cfg['precision']}'. Valid options: fp64, mixed")
    if cfg["orthogonalization"] not in ORTHO:
        raise ConfigError(f"{where}: unknown orthogonalization '{cfg['orthogonalization']}'. Valid options: mgs, cgs2")
    if cfg["orthogonalization"] != "mgs" and typ != "gmres":
        raise ConfigError(f"{where}: orthogonalization applies to gmres")
    if cfg["precision"] == "mixed" and cfg["operator"] != "csr":
        raise ConfigError(f"{where}: mixed precision needs the csr operator")
    if cfg["preconditioner"] not in ("identity", "jacobi"):
        raise ConfigError(f"Unknown preconditioner type: {cfg['preconditioner']}. Valid options: identity, jacobi")
    return cfg


def _monitors(root, base):
    """MonitorConfigParser.cpp (with MonitorTermConfigParser,
    MonitorOutputConfigParser, EvaluationMethodParser, ReductionParser)."""
    mons = []
    for k, m in enumerate(_opt(root, "monitors", None) or []):
        where = f"monitors[{k}]"
        name = str(_req(m, "name", where))
        quantity = _req(m, "quantity", where)
        if quantity != "heat_flux":
            raise ConfigError(f"Unknown monitor quantity: {quantity}")
        method = _opt(m, "flux_method", "weak")
        if method not in ("strong", "weak"):
            raise ConfigError(f"EvaluationMethodParser: unknown method '{method}'. Valid options: strong, weak")
        reduction = _req(m, "reduction", where)
        if reduction not in ("integral", "average"):
            raise ConfigError(f"ReductionParser: unknown reduction '{reduction}'. Valid options: integral, average")
        if "terms" not in m:
            raise ConfigError(f"MonitorConfigParser: monitor '{name}': missing required 'terms' list")
        terms = [(int(_req(t, "boundary", where + ".terms")), float(_opt(t, "coefficient", 1.0)))
                 for t in (m["terms"] or [])]
        out = _opt(m, "output", None) or {}
        f = str(_opt(out, "file", ""))
        mons.append(dict(name=name, method=method, reduction=reduction, terms=terms,
                         console=bool(_opt(out, "console", True)), file=os.path.join(base, f) if f else ""))
    return mons


def parse_heat_config(path):
    import yaml
    with open(path) as f:
        root = yaml.safe_load(f)
    base = os.path.dirname(os.path.abspath(path))
    backend = _opt(root, "backend", "hip")
    if backend != "hip":
        raise ConfigError(f"Unknown backend type: '{backend}'. Valid options: hip")
    cfg = dict(base=base)
    cfg["mesh"] = os.path.join(base, str(_req(_req(root, "mesh", "config"), "file", "mesh")))
    disc = _req(root, "discretization", "config")
    quad = _req(disc, "quadrature", "discretization")
    cfg["quadrature"] = (int(_opt(quad, "xi", 2)), int(_opt(quad, "eta", 2)), int(_opt(quad, "zeta", 2)))
    dof = _opt(disc, "dof_ordering", None)
    if dof and _req(dof, "type", "dof_ordering") not in ("interleaved", "block"):
        raise ConfigError(f"Unknown dof_ordering type: {dof['type']}")
    cfg["arithmetic"] = str(_opt(disc, "arithmetic", "reference"))
    if cfg["arithmetic"] not in ARITHMETIC:
        raise ConfigError(f"discretization: unknown arithmetic '{cfg['arithmetic']}'. Valid options: "
                          + ", ".join(ARITHMETIC))
    # materials (ConductivityConfigParser.cpp, ScalarMaterialPropertyConfigParser.cpp, SpecificHeatConfigParser.cpp)
    mat = _req(root, "materials", "config")
    cond = _req(mat, "conductivity", "materials")
    ctype = _req(cond, "type", "materials.conductivity")
    c = dict(type=ctype, value=1.0, tensor=None, k_expr=None, dk_expr=None)
    if ctype == "constant":
        c["value"] = float(_req(cond, "value", "materials.conductivity"))
    elif ctype == "anisotropic":
        c["tensor"] = [[float(v) for v in row] for row in _req(cond, "tensor", "materials.conductivity")]
    elif ctype in ("temperature_dependent_isotropic", "temperature_dependent_anisotropic"):
        c["k_expr"] = str(_req(cond, "value_expression", "materials.conductivity"))
        c["dk_expr"] = str(_req(cond, "gradient_expression", "materials.conductivity"))
        if ctype.endswith("anisotropic"):
            c["tensor"] = [[float(v) for v in row] for row in _req(cond, "tensor", "materials.conductivity")]
    else:
        raise ConfigError(f"Unknown conductivity type: {ctype}")
    cfg["conductivity"] = c
    cfg["rho"] = float(_req(mat["density"], "value", "materials.density")) if "density" in mat else 1.0
    cp = mat.get("specific_heat")
    if cp is not None and _opt(cp, "type", "constant") != "constant":
        raise ConfigError("materials.specific_heat: only a constant specific heat runs on the device")
    cfg["cp"] = float(_req(cp, "value", "materials.specific_heat")) if cp is not None else 1.0
    # source (SourceConfigParser.cpp)
    cfg["source"] = None
    models = _opt(_opt(root, "physics", {}), "models", {})
    if models and "source" in models:
        src = models["source"]
        if _opt(src, "type", "volumetric") != "volumetric":
            raise ConfigError("physics.models.source: only volumetric sources")
        cfg["source"] = _read_field(src, "physics.models.source", allow_file=True)
    # boundary conditions (BoundaryConditionConfigParser.cpp:26-60): each form
    # registers the condition, value_bc as essential, flux_bc as natural
    # (HeatProblem.tpp:62-112); a flux reads one expression per spatial
    # component
    tags, exprs, natural = [], [], []
    for k, bc in enumerate(_opt(root, "boundary_conditions", []) or []):
        where = f"boundary_conditions[{k}]"
        btype = _req(bc, "type", where)
        if btype not in ("value", "flux"):
            raise ConfigError(f"Unknown boundary type: {btype}")
        tag = int(_req(bc, "boundary", where))
        read = _req(bc, "read", where)
        mode = _opt(read, "mode", "expression")
        for form in _req(bc, "forms", where):
            if form == "value_bc":
                if mode != "expression" or btype != "value":
                    raise ConfigError(f"{where}: boundary values other than expressions are not run on the device")
                tags.append(tag)
                exprs.append(str(_req(read, "expression", where + ".read")))
            elif form == "flux_bc":
                if mode != "expression":
                    raise ConfigError(f"{where}: a flux from a nodal file is not run on the device")
                q = _req(read, "expression", where + ".read")
                if btype != "flux" or not isinstance(q, (list, tuple)):
                    raise ConfigError(f"{where}: a flux_bc reads one expression per spatial component")
                natural.append((tag, [str(e) for e in q]))
            else:
                raise ConfigError(f"Unknown boundary form: {form}")
    cfg["bc_tags"], cfg["bc_exprs"], cfg["natural"] = tags, exprs, natural
    ic = root.get("initial_condition")
    cfg["initial"] = _read_field(ic, "initial_condition", allow_file=True) if ic else ("expression", "0")
    cfg["monitors"] = _monitors(root, base)
    # solver (SolverConfigParser, DriverConfigParser, TimeStepperConfigParser, NonlinearSolverConfigParser)
    solver = _req(root, "solver", "config")
    drv = _req(_req(solver, "driver", "solver"), "type", "solver.driver")
    if drv not in ("steady", "transient"):
        raise ConfigError(f"driver '{drv}' is not run on the device (steady, transient)")
    cfg["driver"] = drv
    cfg["linear"] = _linear(solver["linear"], "solver.linear") if "linear" in solver else None
    nl = solver.get("nonlinear")
    cfg["nonlinear"] = None
    if nl is not None:
        if _req(nl, "type", "solver.nonlinear") != "newton":
            raise ConfigError("solver.nonlinear: only newton runs on the device")
        cfg["nonlinear"] = dict(abs_tol=float(_opt(nl, "absolute_tolerance", 1e-10)),
                                rel_tol=float(_opt(nl, "relative_tolerance", 1e-8)),
                                max_iterations=int(_opt(nl, "max_iterations", 50)),
                                linear=_linear(nl["linear_solver"], "solver.nonlinear.linear_solver")
                                if "linear_solver" in nl else None)
    if cfg["linear"] is None and cfg["nonlinear"] is None:
        raise ConfigError("solver.linear or solver.nonlinear config is required")
    if drv == "transient":
        ts = _req(solver, "timestepper", "solver")
        if _req(ts, "type", "solver.timestepper") != "backward_euler":
            raise ConfigError(f"time stepper '{ts['type']}' is not run on the device (backward_euler)")
        step = _req(ts, "step_size", "solver.timestepper")
        if _req(step, "mode", "solver.timestepper.step_size") != "constant":
            raise ConfigError("solver.timestepper.step_size: only the constant mode runs on the device")
        cfg["t0"] = float(_opt(ts, "t0", 0.0))
        cfg["tf"] = float(_opt(ts, "tf", 0.0))
        cfg["dt"] = float(_req(step, "dt", "solver.timestepper.step_size"))
        if "linear_solver" in ts:
            cfg["linear"] = _linear(ts["linear_solver"], "solver.timestepper.linear_solver")
        if cfg["linear"] is None:
            raise ConfigError("transient runs need a linear solver")
    # output (OutputConfigParser.cpp)
    out = _opt(root, "output", None)
    cfg["output"] = None
    if out is not None:
        fmt = _opt(out, "format", "vtk")
        if fmt not in ("vtk", "vtu"):
            raise ConfigError(f"OutputConfigParser: unknown 'format' value '{fmt}'. Valid options: vtk, vtu")
        cfg["output"] = dict(directory=os.path.join(base, str(_opt(out, "directory", "output"))),
                             write_frequency=int(_opt(out, "write_frequency", 1)),
                             prefix=str(_opt(out, "prefix", "solution")), format=fmt)
    return cfg


# ---------------------------------------------------------------------------
# output: VTK legacy / VTU (ASCII), the element's corners in VTK order
# (VTKWriter::rowMajorToCCW) for Q1 cells, plus the binary nodal field
# ---------------------------------------------------------------------------
_CCW = {4: [0, 1, 3, 2], 8: [0, 1, 3, 2, 4, 5, 7, 6]}
_VTK_TYPE = {4: 9, 8: 12}


def write_vtk_field(path, xyz, ien, npe, dim, nodal, name="T", fmt="vtu"):
    if npe not in _CCW:
        raise ConfigError("VTK output is written for Q1 cells only; the .pndf field is written instead")
    nn = xyz.size // dim
    pts = np.zeros((nn, 3))
    pts[:, :dim] = xyz.reshape(nn, dim)
    cells = ien.reshape(-1, npe)[:, _CCW[npe]]
    ne = cells.shape[0]
    with open(path, "w") as f:
        if fmt == "vtk":
            f.write("# vtk DataFile Version 3.0\nsolver field output\nASCII\nDATASET UNSTRUCTURED_GRID\n")
            f.write(f"POINTS {nn} double\n")
            f.writelines(f"{p[0]:.17g} {p[1]:.17g} {p[2]:.17g}\n" for p in pts)
            f.write(f"CELLS {ne} {ne * (npe + 1)}\n")
            f.writelines(f"{npe} " + " ".join(map(str, c)) + "\n" for c in cells)
            f.write(f"CELL_TYPES {ne}\n")
            f.writelines(f"{_VTK_TYPE[npe]}\n" for _ in range(ne))
            f.write(f"POINT_DATA {nn}\nSCALARS {name} double 1\nLOOKUP_TABLE default\n")
            f.writelines(f"{v:.17g}\n" for v in nodal)
        else:
            f.write('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n')
            f.write(f'<UnstructuredGrid>\n<Piece NumberOfPoints="{nn}" NumberOfCells="{ne}">\n')
            f.write('<Points>\n<DataArray type="Float64" NumberOfComponents="3" format="ascii">\n')
            f.writelines(f"{p[0]:.17g} {p[1]:.17g} {p[2]:.17g}\n" for p in pts)
            f.write('</DataArray>\n</Points>\n<Cells>\n<DataArray type="Int64" Name="connectivity" format="ascii">\n')
            f.writelines(" ".join(map(str, c)) + "\n" for c in cells)
            f.write('</DataArray>\n<DataArray type="Int64" Name="offsets" format="ascii">\n')
            f.write(" ".join(str(npe * (k + 1)) for k in range(ne)) + "\n")
            f.write('</DataArray>\n<DataArray type="UInt8" Name="types" format="ascii">\n')
            f.write(" ".join([str(_VTK_TYPE[npe])] * ne) + "\n")
            f.write(f'</DataArray>\n</Cells>\n<PointData Scalars="{name}">\n')
            f.write(f'<DataArray type="Float64" Name="{name}" NumberOfComponents="1" format="ascii">\n')
            f.writelines(f"{v:.17g}\n" for v in nodal)
            f.write("</DataArray>\n</PointData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")


# ---------------------------------------------------------------------------
# boundary monitors
# ---------------------------------------------------------------------------
class Monitors:
    """The configured monitors of one run (HeatProblem.tpp:133-167): groups
    keyed by (method, reduction) in order of first use, each evaluated once
    per tick over its tags, then every monitor of the group reported."""

    def __init__(self, pb, mons, src_text, transient, quiet):
        self.pb, self.src, self.transient, self.quiet = pb, src_text, transient, quiet
        self.groups = []
        keyed = {}
        bdim = pb.dim - 1
        for m in mons:
            key = (m["method"], m["reduction"])
            if key not in keyed:
                keyed[key] = dict(method=m["method"], reduction=m["reduction"], tags=[], members=[])
                self.groups.append(keyed[key])
            g = keyed[key]
            for tag, _ in m["terms"]:
                if tag not in g["tags"]:
                    g["tags"].append(tag)
            unit = "W" if m["reduction"] == "integral" else f"W/m^{bdim}"  # QuantityUnits.hpp:13-22
            csv = None
            if m["file"]:
                os.makedirs(os.path.dirname(m["file"]) or ".", exist_ok=True)
                csv = open(m["file"], "w")
                csv.write(f"tick,time,{m['name']} [{unit}]\n")
                csv.flush()
            g["members"].append(dict(m, unit=unit, csv=csv))
        self.values = {m["name"]: [] for m in mons}

    def evaluate(self, tick, time, U, Udot=None):
        for g in self.groups:
            if g["method"] == "weak":
                res, meas = self.pb.boundary_reaction(g["tags"], U, Udot=Udot if self.transient else None,
                                                      src=self.src, time=time)
            else:
                res, meas = self.pb.boundary_flux(g["tags"], U, time=time)
            if g["reduction"] == "average":
                res = res / meas
            for m in g["members"]:
                v = 0.0
                for tag, coef in m["terms"]:  # BoundaryQuantityCombination::evaluate
                    v += coef * res[g["tags"].index(tag)]
                self.values[m["name"]].append((tick, time, v))
                if m["console"]:
                    _log(f"monitor '{m['name']}' = {v:.6f} {m['unit']}", self.quiet)
                if m["csv"]:
                    m["csv"].write(f"{float(tick):g},{time:g},{v:g}\n")
                    m["csv"].flush()

    def close(self):
        for g in self.groups:
            for m in g["members"]:
                if m["csv"]:
                    m["csv"].close()


# ---------------------------------------------------------------------------
# the run
# ---------------------------------------------------------------------------
def _log(msg, quiet):
    if not quiet:
        print(msg, flush=True)


def run_heat(config_path, quiet=False, ctx=None):
    """Run one heat-equation configuration on the device; returns a summary
    dict (converged, steps, iterations, the final free vector U as numpy)."""
    cfg = parse_heat_config(config_path)
    mesh = read_pmsh(cfg["mesh"])
    if mesh["basis"] != PMSH_LAGRANGE:
        raise ConfigError("HeatDispatcher: only the Lagrange basis is supported so far")
    dim = mesh["pdim"]
    nq = cfg["quadrature"][:dim]
    if len(set(nq)) != 1:
        raise ConfigError("discretization.quadrature: the device kernels use the same order in every direction")
    _log(f"mesh: {cfg['mesh']} -> {mesh['n_elem']} elements, nsd={mesh['sdim']} npd={dim}", quiet)
    own = ctx is None
    ctx = ctx or Ctx(0)
    try:
        pb = Problem.from_pmsh(ctx, cfg["mesh"], nq[0], cfg["bc_tags"], cfg["bc_exprs"])
        for tag, q in cfg["natural"]:
            if len(q) != mesh["sdim"]:
                raise ConfigError("HeatProblem: flux boundary 'expression' must have SpatialDim components")
        c = cfg["conductivity"]
        tensor = None if c["tensor"] is None else np.array(c["tensor"], np.float64).ravel()
        pb.set_conductivity(c["value"], k_expr=c["k_expr"], dk_expr=c["dk_expr"], tensor=tensor)
        pb.set_mass(cfg["rho"], cfg["cp"])
        pb.set_arithmetic(ARITHMETIC[cfg["arithmetic"]])
        U = pb.vector(0.0)
        mode, val = cfg["initial"]
        if mode == "expression":
            pb.eval_expr(val, cfg.get("t0", 0.0), out=U)
        else:
            nodal, _ = pb.read_pndf(os.path.join(cfg["base"], val))
            _, a2t, _ = pb.dofs()
            U.upload(nodal[a2t])
        src_text, src_nodal = None, None
        if cfg["source"] is not None:
            smode, sval = cfg["source"]
            if smode == "expression":
                src_text = sval
            else:
                nodal, _ = pb.read_pndf(os.path.join(cfg["base"], sval))
                src_nodal = ctx.from_numpy(nodal)
        if src_nodal is not None and any(m["method"] == "weak" for m in cfg["monitors"]):
            raise ConfigError("weak flux monitors with a nodal-file source are not run on the device")
        mon = Monitors(pb, cfg["monitors"], src_text, cfg["driver"] == "transient", quiet) if cfg["monitors"] else None
        try:
            summary = _steady(pb, cfg, U, src_text, src_nodal, quiet, mon) if cfg["driver"] == "steady" else \
                _transient(pb, cfg, U, src_text, src_nodal, mesh, quiet, mon)
        finally:
            if mon:
                mon.close()
        if mon:
            summary["monitors"] = mon.values
        summary["U"] = U.numpy()
        summary["problem"] = dict(n_free=pb.n_free, n_nodes=pb.n_nodes, n_elem=pb.n_elem)
        _log("converged" if summary["converged"] else "did not converge", quiet)
        return summary
    finally:
        if own:
            ctx.close()


def _precond(lin):
    return PC_JACOBI if lin["preconditioner"] == "jacobi" else PC_IDENTITY


@contextlib.contextmanager
def _orthogonalization(ctx, lin):
    """The linear node's GMRES orthogonalisation on the context for one call."""
    ctx.set_gmres_orthogonalization(ORTHO[lin["orthogonalization"]])
    try:
        yield
    finally:
        ctx.set_gmres_orthogonalization(ORTHO_MGS)


def _solve(pb, K, forms, lin, F, U, state=None, time=0.0):
    """One linear solve with the configured operator and solver."""
    if lin["operator"] == "fem":
        if lin["type"] != "cg":
            raise ConfigError("the matrix-free operator runs with cg on the device")
        return pb.cg_fem(F, U, forms, tol=lin["tolerance"], max_iters=lin["max_iterations"], precond=_precond(lin),
                         U=state, time=time)
    mixed = lin["precision"] == "mixed"
    if lin["type"] == "cg":
        if mixed:
            return K.cg_mixed(F, U, tol=lin["tolerance"], max_iters=lin["max_iterations"], precond=_precond(lin))[0]
        return K.cg(F, U, tol=lin["tolerance"], max_iters=lin["max_iterations"], precond=_precond(lin))
    if mixed:
        return K.gmres_mixed(F, U, tol=lin["tolerance"], max_iters=lin["max_iterations"], krylov=lin["krylov_dim"],
                             precond=_precond(lin))[0]
    with _orthogonalization(pb.ctx, lin):
        return K.gmres(F, U, tol=lin["tolerance"], max_iters=lin["max_iterations"], krylov=lin["krylov_dim"],
                       precond=_precond(lin))


def _write(pb, cfg, mesh, U, step, t, quiet):
    out = cfg["output"]
    if out is None or step % max(out["write_frequency"], 1):
        return
    os.makedirs(out["directory"], exist_ok=True)
    stem = os.path.join(out["directory"], f"{out['prefix']}_{step:06d}")
    pb.write_pndf(stem + ".pndf", U, time=t)
    if mesh["npe"] in _CCW:
        write_vtk_field(stem + "." + out["format"], mesh["xyz"], mesh["ien"], mesh["npe"], mesh["sdim"],
                        pb.nodal_field(U, t), fmt=out["format"])
    _log(f"output: {stem}", quiet)


def _steady(pb, cfg, U, src_text, src_nodal, quiet, mon=None):
    mesh = read_pmsh(cfg["mesh"])
    nonlinear = cfg["conductivity"]["k_expr"] is not None
    if nonlinear or cfg["nonlinear"] is not None:
        nl = cfg["nonlinear"] or dict(abs_tol=1e-10, rel_tol=1e-8, max_iterations=50, linear=None)
        lin = nl["linear"] or cfg["linear"]
        if src_nodal is not None:
            raise ConfigError("a nodal-file source with the Newton solver is not run on the device")
        if cfg["natural"]:
            raise ConfigError("flux boundary conditions with the Newton solver are run by hip::HeatProblem, "
                              "not by this command line")
        if lin["operator"] != "csr":
            raise ConfigError("Newton runs with the CSR operator on the device")
        t0 = _time.perf_counter()
        with _orthogonalization(pb.ctx, lin):
            its, res, lin_its = pb.newton(U, src=src_text, abs_tol=nl["abs_tol"], rel_tol=nl["rel_tol"],
                                          max_newton=nl["max_iterations"],
                                          lin_solver={("cg", "fp64"): 0, ("gmres", "fp64"): 1, ("gmres", "mixed"): 2,
                                                      ("cg", "mixed"): 3}[(lin["type"], lin["precision"])],
                                          lin_tol=lin["tolerance"], lin_max=lin["max_iterations"],
                                          krylov=lin["krylov_dim"], precond=_precond(lin))
        conv = its >= 0
        _log(f"newton: {its} iterations, residuals {list(res)} ({_time.perf_counter() - t0:.3f} s)", quiet)
        _write(pb, cfg, mesh, U, 0, 0.0, quiet)
        if mon:
            mon.evaluate(0, 0.0, U)
        return dict(converged=conv, newton_iterations=its, linear_iterations=[int(v) for v in lin_its])
    lin = cfg["linear"]
    forms = [(DIFFUSION, 1.0)]
    F = pb.vector()
    pb.assemble_load(F, src=src_text, nodal=src_nodal)
    if cfg["natural"]:
        pb.apply_natural(F, cfg["natural"])  # SteadyStage::assemble, SteadyStage.hpp:68-70
    pb.apply_essential(F, forms)
    K = None
    if lin["operator"] == "csr":
        K = pb.matrix()
        pb.assemble_matrix(K, forms)
    t0 = _time.perf_counter()
    rep = _solve(pb, K, forms, lin, F, U)
    _log(f"linear: {lin['type']} ({lin['operator']}, {lin['preconditioner']}) {rep.iterations} iterations, "
         f"residual {rep.final_residual:.3e} ({_time.perf_counter() - t0:.3f} s)", quiet)
    _write(pb, cfg, mesh, U, 0, 0.0, quiet)
    if mon:
        mon.evaluate(0, 0.0, U)
    return dict(converged=bool(rep.converged), iterations=int(rep.iterations), final_residual=rep.final_residual)


def _transient(pb, cfg, U, src_text, src_nodal, mesh, quiet, mon=None):
    lin = cfg["linear"]
    t0, tf, dt = cfg["t0"], cfg["tf"], cfg["dt"]
    if cfg["conductivity"]["k_expr"] is not None:
        raise ConfigError("a temperature-dependent conductivity in a transient run is not run on the device")
    _write(pb, cfg, mesh, U, 0, t0, quiet)
    # the rate the weak monitors see: zero at the initial condition, then the
    # one each step's assemble() forms, BackwardEulerStage.hpp:73-74
    Udot = pb.vector(0.0) if mon else None
    if mon:
        mon.evaluate(0, t0, U, Udot)
    start = _time.perf_counter()
    if lin["operator"] == "csr" and lin["type"] == "cg" and src_nodal is None and cfg["output"] is None and not mon \
            and not cfg["natural"]:
        steps, iters = pb.be_run(U, t0, tf, dt, src=src_text, tol=lin["tolerance"], max_iters=lin["max_iterations"],
                                 precond=_precond(lin), mixed=lin["precision"] == "mixed")
        _log(f"backward euler: {steps} steps on the device, {int(np.sum(iters))} CG iterations "
             f"({_time.perf_counter() - start:.3f} s)", quiet)
        return dict(converged=True, steps=int(steps), iterations=[int(v) for v in iters])
    # one device solve per step (BackwardEulerStage.hpp:71-163)
    forms = [(DIFFUSION, 1.0), (MASS | SCALED, 0.0)]
    K = pb.matrix() if lin["operator"] == "csr" else None
    Uprev, F, ms = pb.vector(), pb.vector(), pb.vector()
    t, step, iters = t0, 0, []
    while not (t >= tf - 1e-9):
        h = min(dt, tf - t)
        ct = t + h
        forms[1] = (MASS | SCALED, 1.0 / h)
        pb.ctx.copy(U, Uprev)
        if mon:
            pb.ctx.copy(U, Udot)
            pb.ctx.axpby(-1.0 / h, 1.0 / h, Uprev, Udot)
        pb.assemble_load(F, src=src_text, nodal=src_nodal, time=ct)
        pb.assemble_vector(ms, [(MASS, 1.0)], U=Uprev, selector=SEL_VALUE, time=ct - h)
        pb.ctx.axpy(1.0 / h, ms, F)
        if cfg["natural"]:
            pb.apply_natural(F, cfg["natural"], time=ct)  # BackwardEulerStage.hpp:93
        pb.apply_essential(F, forms, time=ct)
        if K is not None:
            pb.assemble_matrix(K, forms, time=ct)
        rep = _solve(pb, K, forms, lin, F, U, time=ct)
        if not rep.converged:
            raise RsdError(f"backward euler: the linear solve did not converge in step {step}")
        iters.append(int(rep.iterations))
        t, step = ct, step + 1
        _write(pb, cfg, mesh, U, step, t, quiet)
        if mon:
            mon.evaluate(step, t, U, Udot)
    _log(f"backward euler: {step} steps, {sum(iters)} {lin['type']} iterations "
         f"({_time.perf_counter() - start:.3f} s)", quiet)
    return dict(converged=True, steps=step, iterations=iters)


# ---------------------------------------------------------------------------
# the mesh application (src/application/mesh/MeshDispatcher.cpp:18-76): block
# meshes to PMSH
# ---------------------------------------------------------------------------
def run_mesh(config_path, quiet=False):
    import yaml
    with open(config_path) as f:
        root = yaml.safe_load(f)
    base = os.path.dirname(os.path.abspath(config_path))
    m = _req(root, "mesh", "config")
    mtype = _req(m, "type", "mesh")
    if mtype not in ("block2d", "block3d"):
        raise ConfigError(f"mesh type '{mtype}' is not generated here (block2d, block3d)")
    dim = 2 if mtype == "block2d" else 3
    axes = "xyz"[:dim]
    n = [int(_req(m, "n" + a, "mesh")) for a in axes]
    p = [int(_opt(m, "p" + a, 1)) for a in axes]
    ext = []
    for a in axes:
        ext += [float(_req(m, a + "min", "mesh")), float(_req(m, a + "max", "mesh"))]
    out = os.path.join(base, str(_req(_req(root, "output", "config"), "output_file", "output")))
    xyz, ien, rng = block_mesh(dim, n, ext, p)
    npe = int(np.prod([q + 1 for q in p]))
    write_pmsh(out, dict(pdim=dim, sdim=dim, family=PMSH_QUAD if dim == 2 else PMSH_HEX, n_nodes=xyz.size // dim,
                         n_elem=ien.size // npe, npe=npe, faces=2 * dim, orders=p, xyz=xyz, ien=ien, rng=rng))
    _log(f"mesh: {mtype} {n} -> {out}", quiet)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m residuum_hip")
    sub = ap.add_subparsers(dest="app", required=True)
    h = sub.add_parser("heateq", help="run a heat-equation configuration on the device")
    h.add_argument("config")
    h.add_argument("--quiet", action="store_true")
    m = sub.add_parser("mesh", help="write a block mesh as PMSH")
    m.add_argument("config")
    a = ap.parse_args(argv)
    try:
        if a.app == "heateq":
            s = run_heat(a.config, quiet=a.quiet)
            return 0 if s["converged"] else 1
        run_mesh(a.config)
        return 0
    except (ConfigError, RsdError, OSError) as e:
        print(f"error: {e}", file=sys.stderr)
        return 2
