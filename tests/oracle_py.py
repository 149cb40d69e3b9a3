"""ctypes access to the CPU oracle (oracle/liboracle.so) and the reference
harness (oracle/_ref/libref.so).

Test infrastructure only: the oracle is the checker, never the thing
measured or shipped.
"""
import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")

ORDER_REF, ORDER_GPU = 0, 1
DIFFUSION, MASS, TDIFFUSION, TMASS = 1, 2, 3, 4
SCALED = 0x100
SEL_VALUE, SEL_RATE, SEL_OPERAND = 0, 1, 2

# function kinds understood by the reference harness, as expressions
KIND_EXPR = {0: "0", 1: "x", 2: "(1-x)*(1-y)", 3: "10*sin(pi*x)*sin(pi*y)", 4: "x*y",
             5: "x*y*z", 6: "1", 7: "x+2*y+3*z"}

_d = C.POINTER(C.c_double)
_i64 = C.POINTER(C.c_int64)
_i32 = C.POINTER(C.c_int32)


def dp(a):
    return None if a is None else a.ctypes.data_as(_d)


def lp(a):
    return None if a is None else a.ctypes.data_as(_i64)


def ip(a):
    return None if a is None else a.ctypes.data_as(_i32)


class OrcProblem(C.Structure):
    _fields_ = [("dim", C.c_int32), ("p", C.c_int32 * 3), ("nq", C.c_int32 * 3), ("npe", C.c_int32),
                ("n_nodes", C.c_int64), ("n_elem", C.c_int64), ("n_free", C.c_int64),
                ("xyz", _d), ("ien", _i64), ("topo_to_alg", _i32), ("ctag", _i32),
                ("elem_order", _i64), ("kappa_e", _d), ("k_expr", C.c_void_p), ("dk_expr", C.c_void_p),
                ("k_scalar", C.c_double), ("anisotropic", C.c_int32), ("k_tensor", C.c_double * 9),
                ("rho", C.c_double), ("cp", C.c_double), ("cp_expr", C.c_void_p), ("dcp_expr", C.c_void_p),
                ("n_bc", C.c_int32), ("bc_tag", C.c_int32 * 16), ("bc_expr", C.c_void_p * 16)]


class OrcReport(C.Structure):
    _fields_ = [("converged", C.c_int32), ("iterations", C.c_int64), ("initial_residual", C.c_double),
                ("final_residual", C.c_double), ("final_residual_rel", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(ORACLE_SO)
        L.orc_expr_compile.restype = C.c_void_p
        L.orc_expr_compile.argtypes = [C.c_char_p]
        L.orc_expr_free.argtypes = [C.c_void_p]
        L.orc_expr_eval.restype = C.c_double
        L.orc_expr_eval.argtypes = [C.c_void_p] + [C.c_double] * 5
        L.orc_build_constraints.restype = C.c_int64
        L.orc_create_matrix.restype = C.c_int64
        L.orc_colour_order.restype = C.c_int32
        L.orc_dot.restype = C.c_double
        L.orc_be_run.restype = C.c_int64
        L.orc_newton_steady.restype = C.c_int64
        L.orc_adjoint_grad.restype = C.c_double
        _lib = L
    return _lib


def have_ref():
    return os.path.exists(REF_SO)


_ref = None


def ref():
    global _ref
    if _ref is None:
        R = C.CDLL(REF_SO)
        R.ref_create.restype = C.c_void_p
        for name in ["ref_destroy", "ref_set_model", "ref_sizes", "ref_mesh", "ref_dofs", "ref_pattern",
                     "ref_map", "ref_assemble_matrix", "ref_assemble_vector", "ref_load", "ref_lift",
                     "ref_solve", "ref_be_run", "ref_newton"]:
            getattr(R, name).argtypes = None
        R.ref_destroy.argtypes = [C.c_void_p]
        _ref = R
    return _ref


class Expr:
    def __init__(self, text):
        self.text = text
        self.ptr = lib().orc_expr_compile(text.encode())
        if not self.ptr:
            raise ValueError(f"bad expression {text!r}")

    def __call__(self, t=0.0, x=0.0, y=0.0, z=0.0, T=0.0):
        return lib().orc_expr_eval(C.c_void_p(self.ptr), t, x, y, z, T)

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.orc_expr_free(C.c_void_p(self.ptr))


def block_mesh(dim, n, ext, p):
    """BlockMesh2D/3D arrays: (xyz, ien int64, rng int32)."""
    if dim == 2:
        nx, ny = n
        px, py = p
        nn = (nx * px + 1) * (ny * py + 1)
        ne, npe = nx * ny, (px + 1) * (py + 1)
        xyz = np.zeros(nn * 2); ien = np.zeros(ne * npe, np.int64); rng = np.zeros(ne * 4, np.int32)
        lib().orc_block_mesh_2d(C.c_int64(nx), C.c_int64(ny), *[C.c_double(v) for v in ext[:4]],
                                C.c_int64(px), C.c_int64(py), dp(xyz), lp(ien), ip(rng))
    else:
        nx, ny, nz = n
        px, py, pz = p
        nn = (nx * px + 1) * (ny * py + 1) * (nz * pz + 1)
        ne, npe = nx * ny * nz, (px + 1) * (py + 1) * (pz + 1)
        xyz = np.zeros(nn * 3); ien = np.zeros(ne * npe, np.int64); rng = np.zeros(ne * 6, np.int32)
        lib().orc_block_mesh_3d(C.c_int64(nx), C.c_int64(ny), C.c_int64(nz), *[C.c_double(v) for v in ext[:6]],
                                C.c_int64(px), C.c_int64(py), C.c_int64(pz), dp(xyz), lp(ien), ip(rng))
    return xyz, ien, rng


# flux vector kinds of the reference harness (oracle/ref_harness.cpp FluxKindFn)
FLUX_KINDS = [("1.5", "-0.5", "0.25"), ("x", "y*y", "x*z"), ("sin(pi*x)+t", "cos(y)", "x*y")]


@dataclass
class Case:
    """A discretised heat problem on a block mesh, with every oracle array."""
    dim: int
    n: tuple
    p: tuple
    nq: int
    tags: list
    bc_exprs: list
    ext: tuple = None
    kappa: np.ndarray = None
    k_scalar: float = 1.0
    k_expr: str = None
    dk_expr: str = None
    tensor: tuple = None
    rho: float = 1.0
    cp: float = 1.0
    order: int = ORDER_REF
    _keep: list = field(default_factory=list)

    def __post_init__(self):
        if self.ext is None:
            self.ext = (0.0, 1.0) * self.dim
        self.xyz, self.ien, self.rng = block_mesh(self.dim, self.n, self.ext, self.p)
        self.npe = int(np.prod([q + 1 for q in self.p]))
        self.n_nodes = self.xyz.size // self.dim
        self.n_elem = self.ien.size // self.npe
        p3 = np.array(list(self.p) + [1] * (3 - self.dim), np.int32)
        self.t2a = np.zeros(self.n_nodes, np.int32)
        self.a2t = np.zeros(self.n_nodes, np.int32)
        self.ctag = np.zeros(self.n_nodes, np.int32)
        tags = np.array(self.tags, np.int32)
        self.n_free = lib().orc_build_constraints(self.dim, ip(p3), C.c_int64(self.n_nodes), C.c_int64(self.n_elem),
                                                  lp(self.ien), ip(self.rng), len(tags), ip(tags),
                                                  ip(self.t2a), ip(self.a2t), ip(self.ctag))
        self.a2t = self.a2t[:self.n_free].copy()
        self.rp = np.zeros(self.n_free + 1, np.int64)
        self.nnz = lib().orc_create_matrix(C.c_int64(self.n_elem), self.npe, lp(self.ien), ip(self.t2a),
                                           C.c_int64(self.n_free), lp(self.rp), None)
        self.ci = np.zeros(self.nnz, np.int64)
        lib().orc_create_matrix(C.c_int64(self.n_elem), self.npe, lp(self.ien), ip(self.t2a),
                                C.c_int64(self.n_free), lp(self.rp), lp(self.ci))
        self.map = np.zeros(self.n_elem * self.npe * self.npe, np.int64)
        lib().orc_elem_nnz_map(C.c_int64(self.n_elem), self.npe, lp(self.ien), ip(self.t2a),
                               lp(self.rp), lp(self.ci), lp(self.map))
        self.colour = np.zeros(self.n_elem, np.int32)
        self.elem_order = np.zeros(self.n_elem, np.int64)
        self.n_colours = lib().orc_colour_order(C.c_int64(self.n_elem), self.npe, lp(self.ien),
                                                C.c_int64(self.n_nodes), ip(self.colour), lp(self.elem_order))
        self._exprs = [Expr(e) for e in self.bc_exprs]
        self._k = Expr(self.k_expr) if self.k_expr else None
        self._dk = Expr(self.dk_expr) if self.dk_expr else None
        self.P = self.make_problem(self.order)

    def make_problem(self, order):
        P = OrcProblem()
        P.dim = self.dim
        for i in range(3):
            P.p[i] = self.p[i] if i < self.dim else 0
            P.nq[i] = self.nq if i < self.dim else 0
        P.npe = self.npe
        P.n_nodes, P.n_elem, P.n_free = self.n_nodes, self.n_elem, self.n_free
        P.xyz, P.ien, P.topo_to_alg, P.ctag = dp(self.xyz), lp(self.ien), ip(self.t2a), ip(self.ctag)
        P.elem_order = lp(self.elem_order) if order == ORDER_GPU else None
        P.kappa_e = dp(self.kappa) if self.kappa is not None else None
        P.k_expr = self._k.ptr if self._k else None
        P.dk_expr = self._dk.ptr if self._dk else None
        P.k_scalar = self.k_scalar
        if self.tensor is not None:
            P.anisotropic = 1
            for i, v in enumerate(self.tensor):
                P.k_tensor[i] = v
        P.rho, P.cp = self.rho, self.cp
        P.n_bc = len(self.tags)
        for i, (t, e) in enumerate(zip(self.tags, self._exprs)):
            P.bc_tag[i] = t
            P.bc_expr[i] = e.ptr
        return P

    def set_order(self, order):
        self.order = order
        self.P = self.make_problem(order)

    # -- assembly -------------------------------------------------------
    def assemble_matrix(self, forms, U=None, Udot=None, time=0.0):
        kinds = np.array([f[0] for f in forms], np.int32)
        coefs = np.array([f[1] for f in forms], np.float64)
        val = np.zeros(self.nnz)
        lib().orc_assemble_matrix(C.byref(self.P), C.c_double(time), len(forms), ip(kinds), dp(coefs),
                                  dp(U), dp(Udot), lp(self.map), C.c_int64(self.nnz), dp(val))
        return val

    def assemble_vector(self, forms, U=None, Udot=None, operand=None, selector=SEL_VALUE, time=0.0):
        kinds = np.array([f[0] for f in forms], np.int32)
        coefs = np.array([f[1] for f in forms], np.float64)
        F = np.zeros(self.n_free)
        lib().orc_assemble_vector(C.byref(self.P), C.c_double(time), len(forms), ip(kinds), dp(coefs),
                                  dp(U), dp(Udot), dp(operand), selector, dp(F))
        return F

    def load(self, src=None, nodal=None, time=0.0):
        F = np.zeros(self.n_free)
        e = Expr(src) if src else None
        lib().orc_assemble_load(C.byref(self.P), C.c_double(time), C.c_void_p(e.ptr) if e else None,
                                dp(nodal), dp(F))
        return F

    def lift(self, forms, F, time=0.0):
        kinds = np.array([f[0] for f in forms], np.int32)
        coefs = np.array([f[1] for f in forms], np.float64)
        F = F.copy()
        lib().orc_apply_essential(C.byref(self.P), C.c_double(time), len(forms), ip(kinds), dp(coefs), dp(F))
        return F

    def boundary_flux(self, tags, U, time=0.0):
        """Strong boundary heat flux per tag (integral, face measure)."""
        t = np.array(tags, np.int32)
        res, meas = np.zeros(len(t)), np.zeros(len(t))
        lib().orc_boundary_flux(C.byref(self.P), C.c_double(time), dp(U), ip(self.rng), len(t), ip(t), dp(res),
                                dp(meas))
        return res, meas

    def apply_natural(self, entries, F, time=0.0):
        """F += natural (flux) BC terms; entries: list of (tag, flux) with
        flux a kind number (FLUX_KINDS expressions) or a nodal array
        [n_nodes * dim]."""
        d = self.dim
        n = len(entries)
        tags = np.array([t for t, _ in entries], np.int32)
        keep, ex, nod = [], (C.c_void_p * max(n * d, 1))(), (C.c_void_p * max(n, 1))()
        for k, (_, fl) in enumerate(entries):
            if isinstance(fl, (int, np.integer)):
                for i in range(d):
                    e = Expr(FLUX_KINDS[fl][i])
                    keep.append(e)
                    ex[k * d + i] = e.ptr
            else:
                a = np.ascontiguousarray(fl, np.float64)
                keep.append(a)
                nod[k] = a.ctypes.data
        out = np.array(F, np.float64)
        lib().orc_apply_natural(C.byref(self.P), C.c_double(time), ip(self.rng), C.c_int32(n), ip(tags), ex, nod,
                                dp(out))
        return out

    def boundary_reaction(self, tags, U, Udot=None, src=None, time=0.0):
        """Weak boundary reaction per tag (sum, face measure), ascending
        element order."""
        t = np.array(tags, np.int32)
        res, meas = np.zeros(len(t)), np.zeros(len(t))
        e = Expr(src) if src else None
        lib().orc_boundary_reaction(C.byref(self.P), C.c_double(time), C.c_void_p(e.ptr) if e else None, dp(U),
                                    dp(Udot), ip(self.rng), len(t), ip(t), dp(res), dp(meas))
        return res, meas

    def assemble_diagonal(self, forms, U=None, time=0.0):
        kinds = np.array([f[0] for f in forms], np.int32)
        coefs = np.array([f[1] for f in forms], np.float64)
        d = np.zeros(self.n_free)
        lib().orc_assemble_diagonal(C.byref(self.P), C.c_double(time), len(forms), ip(kinds), dp(coefs), dp(U), dp(d))
        return d

    # -- solvers --------------------------------------------------------
    def cg_fem(self, forms, b, x0=None, tol=1e-10, max_iters=1000, precond=1, U=None, Udot=None, time=0.0,
               order=None, hist=False):
        kinds = np.array([f[0] for f in forms], np.int32)
        coefs = np.array([f[1] for f in forms], np.float64)
        x = np.zeros(self.n_free) if x0 is None else x0.astype(np.float64).copy()
        rep = OrcReport()
        h = np.zeros(max_iters + 1) if hist else None
        lib().orc_cg_fem(C.byref(self.P), C.c_double(time), len(forms), ip(kinds), dp(coefs), dp(U), dp(Udot),
                         dp(b), dp(x), C.c_double(tol), C.c_int64(max_iters), precond,
                         self.order if order is None else order, C.byref(rep), dp(h))
        return (x, rep, h) if hist else (x, rep)

    def spmv(self, val, x):
        y = np.zeros(self.n_free)
        lib().orc_spmv(C.c_int64(self.n_free), lp(self.rp), lp(self.ci), dp(val), dp(x), dp(y))
        return y

    def cg(self, val, b, x0=None, tol=1e-10, max_iters=1000, precond=1, order=None, hist=False):
        x = np.zeros(self.n_free) if x0 is None else x0.astype(np.float64).copy()
        rep = OrcReport()
        h = np.zeros(max_iters + 1) if hist else None
        lib().orc_cg(C.c_int64(self.n_free), lp(self.rp), lp(self.ci), dp(val), dp(b), dp(x), C.c_double(tol),
                     C.c_int64(max_iters), precond, self.order if order is None else order, C.byref(rep), dp(h))
        return (x, rep, h) if hist else (x, rep)

    def gmres(self, val, b, x0=None, tol=1e-10, max_iters=1000, krylov=30, precond=1, order=None):
        x = np.zeros(self.n_free) if x0 is None else x0.astype(np.float64).copy()
        rep = OrcReport()
        lib().orc_gmres(C.c_int64(self.n_free), lp(self.rp), lp(self.ci), dp(val), dp(b), dp(x), C.c_double(tol),
                        C.c_int64(max_iters), C.c_int64(krylov), precond, self.order if order is None else order,
                        C.byref(rep))
        return x, rep

    def be_run(self, t0, tf, dt, U0=None, src=None, cg_tol=1e-8, cg_max=1000, precond=1, n_steps_max=100000,
               hist=False):
        U = np.zeros(self.n_free) if U0 is None else U0.astype(np.float64).copy()
        nmax = min(n_steps_max, int(round((tf - t0) / dt)) + 2)
        iters = np.zeros(nmax, np.int64)
        Uh = np.zeros(nmax * self.n_free) if hist else None
        e = Expr(src) if src else None
        steps = lib().orc_be_run(C.byref(self.P), C.c_double(t0), C.c_double(tf), C.c_double(dt),
                                 C.c_void_p(e.ptr) if e else None, lp(self.rp), lp(self.ci), lp(self.map),
                                 C.c_int64(self.nnz), C.c_double(cg_tol), C.c_int64(cg_max), precond, self.order,
                                 dp(U), dp(Uh), C.c_int64(nmax), lp(iters))
        out = {"U": U, "steps": steps, "iters": iters[:max(steps, 0)]}
        if hist:
            out["U_hist"] = Uh[:max(steps, 0) * self.n_free].reshape(-1, self.n_free)
        return out

    def steady(self, src=None, nodal=None, solver=0, tol=1e-10, max_iters=1000, krylov=30, precond=1, U0=None):
        U = np.zeros(self.n_free) if U0 is None else U0.astype(np.float64).copy()
        rep = OrcReport()
        e = Expr(src) if src else None
        lib().orc_steady_linear(C.byref(self.P), C.c_void_p(e.ptr) if e else None, dp(nodal), lp(self.rp),
                                lp(self.ci), lp(self.map), C.c_int64(self.nnz), solver, C.c_double(tol),
                                C.c_int64(max_iters), C.c_int64(krylov), precond, self.order, dp(U), C.byref(rep))
        return U, rep

    def newton(self, src=None, abs_tol=1e-10, rel_tol=1e-10, max_newton=50, lin_solver=1, lin_tol=1e-12,
               lin_max=5000, krylov=30, precond=1, U0=None):
        U = np.zeros(self.n_free) if U0 is None else U0.astype(np.float64).copy()
        rh = np.zeros(max_newton)
        li = np.zeros(max_newton, np.int64)
        e = Expr(src) if src else None
        its = lib().orc_newton_steady(C.byref(self.P), C.c_void_p(e.ptr) if e else None, lp(self.rp), lp(self.ci),
                                      lp(self.map), C.c_int64(self.nnz), C.c_double(abs_tol), C.c_double(rel_tol),
                                      C.c_int64(max_newton), lin_solver, C.c_double(lin_tol), C.c_int64(lin_max),
                                      C.c_int64(krylov), precond, self.order, dp(U), dp(rh), lp(li))
        n = max(its, 0) + 1
        return U, its, rh[:n], li[:max(its, 0)]

    def adjoint(self, t0, dt, n_steps, obs, t_obs, U0=None, cg_tol=1e-13, cg_max=5000):
        U = np.zeros(self.n_free) if U0 is None else U0
        grad = np.zeros(self.n_elem)
        UN = np.zeros(self.n_free)
        loss = lib().orc_adjoint_grad(C.byref(self.P), C.c_double(t0), C.c_double(dt), C.c_int64(n_steps),
                                      lp(self.rp), lp(self.ci), lp(self.map), C.c_int64(self.nnz),
                                      C.c_double(cg_tol), C.c_int64(cg_max), self.order, dp(U),
                                      C.c_int64(len(obs)), lp(np.asarray(obs, np.int64)),
                                      dp(np.asarray(t_obs, np.float64)), dp(grad), dp(UN))
        return loss, grad, UN

    def nodal_full(self, U, time=0.0):
        """Node-indexed temperature (free values + Dirichlet values)."""
        out = np.zeros(self.n_nodes)
        for nd in range(self.n_nodes):
            a = self.t2a[nd]
            if a >= 0:
                out[nd] = U[a]
            else:
                x = self.xyz[nd * self.dim:(nd + 1) * self.dim]
                v = 0.0
                for t, e in zip(self.tags, self._exprs):
                    if t == self.ctag[nd]:
                        v = e(time, *x, *([0.0] * (3 - self.dim)))
                out[nd] = v
        return out


def dot(x, y, order):
    return lib().orc_dot(C.c_int64(x.size), dp(x), dp(y), order)


def normal_field(seed, stream, n):
    out = np.zeros(n)
    lib().orc_normal_field(C.c_uint64(seed), C.c_uint32(stream), C.c_int64(n), dp(out))
    return out


def uniform_field(seed, stream, n):
    out = np.zeros(n)
    lib().orc_uniform_field(C.c_uint64(seed), C.c_uint32(stream), C.c_int64(n), dp(out))
    return out


class RefCase:
    """The same problem driven through the reference's own headers."""

    def __init__(self, dim, n, p, nq, tags, kinds, ext=None):
        R = ref()
        ext = ext if ext is not None else (0.0, 1.0) * dim
        n3 = list(n) + [1] * (3 - dim)
        p3 = list(p) + [1] * (3 - dim)
        e6 = np.array(list(ext) + [0.0] * (6 - len(ext)), np.float64)
        t = np.array(tags, np.int32)
        k = np.array(kinds, np.int32)
        self.h = C.c_void_p(R.ref_create(C.c_int(dim), C.c_int64(n3[0]), C.c_int64(n3[1]), C.c_int64(n3[2]), dp(e6),
                                         C.c_int64(p3[0]), C.c_int64(p3[1]), C.c_int64(p3[2]), C.c_int(nq),
                                         C.c_int(len(t)), ip(t), ip(k)))
        s = np.zeros(5, np.int64)
        R.ref_sizes(self.h, lp(s))
        self.n_nodes, self.n_elem, self.npe, self.n_free, self.nnz = [int(v) for v in s]
        self.dim = dim

    def __del__(self):
        if _ref is not None and getattr(self, "h", None):
            _ref.ref_destroy(self.h)

    def set_model(self, tdep=0, ka=1.0, kb=0.0, tensor=None, rho=1.0, cp=1.0):
        tz = np.array(tensor if tensor is not None else [1, 0, 0, 0, 1, 0, 0, 0, 1], np.float64)
        ref().ref_set_model(self.h, C.c_int(tdep), C.c_double(ka), C.c_double(kb), C.c_int(tensor is not None),
                            dp(tz), C.c_double(rho), C.c_double(cp))

    def mesh(self):
        xyz = np.zeros(self.n_nodes * self.dim); ien = np.zeros(self.n_elem * self.npe, np.int64)
        rng = np.zeros(self.n_elem * 2 * self.dim, np.int32)
        ref().ref_mesh(self.h, dp(xyz), lp(ien), ip(rng))
        return xyz, ien, rng

    def dofs(self):
        t2a = np.zeros(self.n_nodes, np.int32); ct = np.zeros(self.n_nodes, np.int32)
        ref().ref_dofs(self.h, ip(t2a), ip(ct))
        return t2a, ct

    def pattern(self):
        rp = np.zeros(self.n_free + 1, np.int64); ci = np.zeros(self.nnz, np.int64)
        ref().ref_pattern(self.h, lp(rp), lp(ci))
        return rp, ci

    def map(self):
        m = np.zeros(self.n_elem * self.npe * self.npe, np.int64)
        ref().ref_map(self.h, lp(m))
        return m

    def assemble_matrix(self, kind_set, coef=1.0, U=None, kappa=None):
        val = np.zeros(self.nnz)
        ref().ref_assemble_matrix(self.h, C.c_int(kind_set), C.c_double(coef), dp(U), dp(kappa), dp(val))
        return val

    def assemble_vector(self, kind_set, U=None, kappa=None):
        F = np.zeros(self.n_free)
        ref().ref_assemble_vector(self.h, C.c_int(kind_set), dp(U), dp(kappa), dp(F))
        return F

    def load(self, kind=0, nodal=None):
        F = np.zeros(self.n_free)
        ref().ref_load(self.h, C.c_int(kind), dp(nodal), dp(F))
        return F

    def lift(self, kind_set, F, coef=1.0, kappa=None):
        F = F.copy()
        ref().ref_lift(self.h, C.c_int(kind_set), C.c_double(coef), dp(kappa), dp(F))
        return F

    def boundary_flux(self, tags, U, time=0.0):
        t = np.array(tags, np.int32)
        res, meas = np.zeros(len(t)), np.zeros(len(t))
        ref().ref_boundary_flux(self.h, C.c_double(time), dp(U), C.c_int(len(t)), ip(t), dp(res), dp(meas))
        return res, meas

    def natural(self, entries, F, time=0.0):
        n = len(entries)
        tags = np.array([t for t, _ in entries], np.int32)
        kinds = np.array([fl if isinstance(fl, (int, np.integer)) else 0 for _, fl in entries], np.int32)
        keep, nod = [], (C.c_void_p * max(n, 1))()
        for k, (_, fl) in enumerate(entries):
            if not isinstance(fl, (int, np.integer)):
                a = np.ascontiguousarray(fl, np.float64)
                keep.append(a)
                nod[k] = a.ctypes.data
        out = np.array(F, np.float64)
        ref().ref_natural(self.h, C.c_double(time), C.c_int(n), ip(tags), ip(kinds), nod, dp(out))
        return out

    def boundary_reaction(self, tags, U, Udot=None, src_kind=0, time=0.0):
        t = np.array(tags, np.int32)
        res, meas = np.zeros(len(t)), np.zeros(len(t))
        ref().ref_boundary_reaction(self.h, C.c_double(time), C.c_int(src_kind), dp(U), dp(Udot), None,
                                    C.c_int(len(t)), ip(t), dp(res), dp(meas))
        return res, meas

    def assemble_diagonal(self, kind_set, coef=1.0, U=None, kappa=None):
        d = np.zeros(self.n_free)
        ref().ref_assemble_diagonal(self.h, C.c_int(kind_set), C.c_double(coef), dp(U), dp(kappa), dp(d))
        return d

    def cg_fem(self, kind_set, b, coef=1.0, U=None, kappa=None, precond=1, x0=None, tol=1e-10, max_iters=1000):
        x = np.zeros(self.n_free) if x0 is None else x0.astype(np.float64).copy()
        it = np.zeros(1, np.int64)
        ok = ref().ref_cg_fem(self.h, C.c_int(kind_set), C.c_double(coef), dp(U), dp(kappa), C.c_int(precond), dp(b),
                              dp(x), C.c_double(tol), C.c_int64(max_iters), lp(it))
        return x, bool(ok), int(it[0])

    def solve(self, solver, precond, val, b, x0=None, tol=1e-10, max_iters=1000, krylov=30):
        x = np.zeros(self.n_free) if x0 is None else x0.astype(np.float64).copy()
        it = np.zeros(1, np.int64)
        ok = ref().ref_solve(self.h, C.c_int(solver), C.c_int(precond), dp(val), dp(b), dp(x), C.c_double(tol),
                             C.c_int64(max_iters), C.c_int64(krylov), lp(it))
        return x, bool(ok), int(it[0])

    def be_run(self, t0, tf, dt, U0=None, tol=1e-8, max_iters=1000, kappa=None):
        U = np.zeros(self.n_free) if U0 is None else U0.astype(np.float64).copy()
        iters = np.zeros(int(round((tf - t0) / dt)) + 2, np.int64)
        steps = ref().ref_be_run(self.h, C.c_double(t0), C.c_double(tf), C.c_double(dt), C.c_double(tol),
                                 C.c_int64(max_iters), dp(kappa), dp(U), lp(iters))
        return U, steps, iters[:max(steps, 0)]

    def newton(self, src_kind=0, abs_tol=1e-10, rel_tol=1e-10, max_newton=50, lin_tol=1e-12, lin_max=5000,
               krylov=30, U0=None):
        U = np.zeros(self.n_free) if U0 is None else U0.astype(np.float64).copy()
        li = np.zeros(max_newton, np.int64)
        its = ref().ref_newton(self.h, C.c_double(abs_tol), C.c_double(rel_tol), C.c_int64(max_newton),
                               C.c_int(src_kind), C.c_double(lin_tol), C.c_int64(lin_max), C.c_int64(krylov),
                               dp(U), lp(li))
        return U, its, li[:max(its, 0)]
