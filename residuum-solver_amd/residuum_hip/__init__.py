"""Python access to libresiduum_hip.so, the MI355X (gfx950) backend for
Residuum's heat-equation hot path.

A thin ctypes layer over the C-ABI declared in include/residuum_hip.h.
There is no fallback: if the shared library is missing or no GPU is
visible, the calls raise.
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

from . import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libresiduum_hip.so")

OK = 0
COPY_H2H, COPY_H2D, COPY_D2H, COPY_D2D = 0, 1, 2, 3
DIFFUSION, MASS, TDIFFUSION, TMASS = 1, 2, 3, 4
SCALED = 0x100
SEL_VALUE, SEL_RATE, SEL_OPERAND = 0, 1, 2
PC_IDENTITY, PC_JACOBI = 0, 1
TOL_RELATIVE, TOL_ABSOLUTE = 0, 1
ARITH_REFERENCE, ARITH_CONTRACTED, ARITH_UNIFORM, ARITH_FACTORED = 0, 1, 2, 3
ORTHO_MGS, ORTHO_CGS2 = 0, 1

_d = C.POINTER(C.c_double)
_f = C.POINTER(C.c_float)
_i64 = C.POINTER(C.c_int64)
_i32 = C.POINTER(C.c_int32)
_vp = C.c_void_p


class RsdError(RuntimeError):
    pass


class Report(C.Structure):
    _fields_ = [("converged", C.c_int32), ("iterations", C.c_int64), ("initial_residual", C.c_double),
                ("final_residual", C.c_double), ("final_residual_rel", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RsdError(f"{LIB_PATH} not built: run __graft_entry__.build() or make -C residuum-solver_amd")
        L = C.CDLL(LIB_PATH)
        try:
            _abi.bind(L)
        except RuntimeError as e:
            raise RsdError(f"{LIB_PATH}: {e}") from None
        _lib = L
    return _lib


def check(rc):
    if rc != OK:
        msg = lib().rsd_last_error_string().decode()
        raise RsdError(f"residuum_hip error {rc}: {msg}")
    return rc


def dptr(a):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return _vp(a.ptr)
    if hasattr(a, "data_ptr"):
        return _vp(a.data_ptr())
    return _vp(a)


# Contexts still open at interpreter exit are closed by an atexit hook, while
# the HIP runtime is still up: objects left to module teardown would
# otherwise free device memory after the runtime's own destructors ran.
_live_contexts = weakref.WeakSet()


@atexit.register
def _close_live_contexts():
    for c in list(_live_contexts):
        try:
            c.close()
        except Exception:
            pass


class Ctx:
    """rsd_ctx: one HIP stream on a device."""

    def __init__(self, device=0):
        self.h = _vp()
        self._children = weakref.WeakSet()
        check(lib().rsd_ctx_create(C.c_int(device), C.byref(self.h)))
        _live_contexts.add(self)

    def _adopt(self, obj):
        self._children.add(obj)

    def sync(self):
        check(lib().rsd_ctx_sync(self.h))

    def set_gmres_orthogonalization(self, mode):
        """ORTHO_MGS (default, the reference's modified Gram-Schmidt, bitwise)
        or ORTHO_CGS2 (classical Gram-Schmidt, reorthogonalised once: three
        folds per Arnoldi step) for this context's later GMRES solves (FP64,
        and the mixed solver's FP32 inner GMRES)."""
        check(lib().rsd_ctx_set_gmres_orthogonalization(self.h, C.c_int32(mode)))

    def close(self):
        """Release every object created on this context, then the context."""
        if self.h:
            kids = list(self._children)
            order = {"Batch": 0, "Matrix": 1, "DeviceArray": 2, "Problem": 3}
            for k in sorted(kids, key=lambda o: order.get(type(o).__name__, 9)):
                k.close()
            lib().rsd_ctx_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- memory -----------------------------------------------------------
    def empty(self, n, dtype=np.float64):
        return DeviceArray(self, n, dtype)

    def from_numpy(self, a):
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.size, a.dtype)
        d.upload(a)
        return d

    def zeros(self, n, dtype=np.float64):
        d = DeviceArray(self, n, dtype)
        if n:
            check(lib().rsd_memset_zero(self.h, _vp(d.ptr), C.c_size_t(d.nbytes)))
        return d

    # -- BLAS-1 (FP64 or FP32 device arrays; FP32 dot / norm2 sum in FP64) --
    @staticmethod
    def _kind(*arrs):
        def f32(a):
            return str(a.dtype) in ("float32", "torch.float32")

        if any(f32(a) != f32(arrs[0]) for a in arrs):
            raise TypeError("BLAS-1: mixed FP32 / FP64 operands")
        return ("f32", C.c_float) if f32(arrs[0]) else ("f64", C.c_double)

    def dot(self, x, y):
        k, _ = self._kind(x, y)
        out = C.c_double()
        check(getattr(lib(), "rsd_dot_" + k)(self.h, C.c_int64(x.n), dptr(x), dptr(y), None, C.byref(out)))
        return out.value

    def norm2(self, x):
        k, _ = self._kind(x)
        out = C.c_double()
        check(getattr(lib(), "rsd_norm2_" + k)(self.h, C.c_int64(x.n), dptr(x), None, C.byref(out)))
        return out.value

    def axpy(self, a, x, y):
        k, sc = self._kind(x, y)
        check(getattr(lib(), "rsd_axpy_" + k)(self.h, C.c_int64(x.n), sc(a), dptr(x), dptr(y)))

    def axpby(self, a, b, x, y):
        k, sc = self._kind(x, y)
        check(getattr(lib(), "rsd_axpby_" + k)(self.h, C.c_int64(x.n), sc(a), sc(b), dptr(x), dptr(y)))

    def scal(self, a, x):
        k, sc = self._kind(x)
        check(getattr(lib(), "rsd_scal_" + k)(self.h, C.c_int64(x.n), sc(a), dptr(x)))

    def copy(self, x, y):
        k, _ = self._kind(x, y)
        check(getattr(lib(), "rsd_copy_" + k)(self.h, C.c_int64(x.n), dptr(x), dptr(y)))

    def mul(self, x, y, z):
        k, _ = self._kind(x, y, z)
        check(getattr(lib(), "rsd_mul_" + k)(self.h, C.c_int64(x.n), dptr(x), dptr(y), dptr(z)))

    def fill(self, x, v):
        if x.dtype == np.float32:
            check(lib().rsd_fill_f32(self.h, dptr(x), C.c_int64(x.n), C.c_float(v)))
        else:
            check(lib().rsd_fill_f64(self.h, dptr(x), C.c_int64(x.n), C.c_double(v)))


class DeviceArray:
    """A device buffer allocated through rsd_malloc."""

    def __init__(self, ctx, n, dtype=np.float64):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = _vp()
        check(lib().rsd_malloc(ctx.h, C.c_size_t(self.nbytes), C.byref(p)))
        self.ptr = p.value or 0
        ctx._adopt(self)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        if self.n:
            check(lib().rsd_memcpy(self.ctx.h, _vp(self.ptr), a.ctypes.data_as(_vp), C.c_size_t(self.nbytes), COPY_H2D))

    def numpy(self):
        out = np.empty(self.n, self.dtype)
        if self.n:
            check(lib().rsd_memcpy(self.ctx.h, out.ctypes.data_as(_vp), _vp(self.ptr), C.c_size_t(self.nbytes), COPY_D2H))
        return out

    def free(self):
        if self.ptr and self.ctx.h:
            lib().rsd_free(self.ctx.h, _vp(self.ptr))
        self.ptr = 0

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _cstrs(items):
    arr = (C.c_char_p * max(1, len(items)))()
    for i, s in enumerate(items):
        arr[i] = s.encode()
    return arr


def _forms(forms):
    kinds = (C.c_int32 * max(1, len(forms)))(*[int(f[0]) for f in forms])
    coefs = (C.c_double * max(1, len(forms)))(*[float(f[1]) for f in forms])
    return len(forms), kinds, coefs


PMSH_QUAD, PMSH_HEX = 0, 2  # mesh/ElementFamily.hpp
PMSH_LAGRANGE = 0           # mesh/BasisType.hpp


def block_mesh(dim, n, ext, p):
    """BlockMesh2D/3D::generate on the host: (xyz, ien int64, rng int32)."""
    n = list(n) + [1] * (3 - dim)
    p = list(p) + [p[0]] * (3 - dim)
    P = p[0]
    nn = int(np.prod([n[i] * P + 1 for i in range(dim)]))
    ne = int(np.prod(n[:dim]))
    npe = (P + 1) ** dim
    xyz = np.zeros(nn * dim)
    ien = np.zeros(ne * npe, np.int64)
    rng = np.zeros(ne * 2 * dim, np.int32)
    e6 = (C.c_double * 6)(*(list(ext) + [0.0] * (6 - len(ext))))
    check(lib().rsd_block_mesh(C.c_int32(dim), (C.c_int64 * 3)(*n), e6, (C.c_int32 * 3)(*p),
                               xyz.ctypes.data_as(_d), ien.ctypes.data_as(_i64), rng.ctypes.data_as(_i32)))
    return xyz, ien, rng


def read_pmsh(path):
    """MeshIO::readBinary: a PMSH v1 file as a dict of header fields and the
    xyz / ien / rng arrays."""
    info = (C.c_int64 * 12)()
    check(lib().rsd_pmsh_info(os.fsencode(path), info))
    pdim, sdim, fam, basis, nn, ne, npe, faces, nord = [int(v) for v in info[:9]]
    xyz = np.zeros(nn * sdim)
    ien = np.zeros(ne * npe, np.int64)
    rng = np.zeros(ne * faces, np.int32)
    check(lib().rsd_pmsh_read(os.fsencode(path), xyz.ctypes.data_as(_d), ien.ctypes.data_as(_i64),
                              rng.ctypes.data_as(_i32)))
    return dict(pdim=pdim, sdim=sdim, family=fam, basis=basis, n_nodes=nn, n_elem=ne, npe=npe, faces=faces,
                orders=[int(v) for v in info[9:9 + min(nord, 3)]], xyz=xyz, ien=ien, rng=rng)


def write_pmsh(path, mesh):
    """MeshIO::writeBinary of a dict shaped as read_pmsh returns."""
    xyz = np.ascontiguousarray(mesh["xyz"], np.float64)
    ien = np.ascontiguousarray(mesh["ien"], np.int64)
    rng = np.ascontiguousarray(mesh["rng"], np.int32)
    orders = (C.c_int32 * max(1, len(mesh["orders"])))(*mesh["orders"])
    check(lib().rsd_pmsh_write(os.fsencode(path), C.c_int32(mesh["pdim"]), C.c_int32(mesh["sdim"]),
                               C.c_int32(mesh["family"]), C.c_int32(mesh.get("basis", PMSH_LAGRANGE)),
                               C.c_int64(mesh["n_nodes"]), C.c_int64(mesh["n_elem"]), C.c_int32(mesh["npe"]),
                               C.c_int32(mesh["faces"]), C.c_int32(len(mesh["orders"])), orders,
                               xyz.ctypes.data_as(_d), ien.ctypes.data_as(_i64), rng.ctypes.data_as(_i32)))


class Problem:
    """rsd_problem: mesh + DOFs + pattern + element map + colouring in HBM."""

    def __init__(self, ctx, dim, n, p, nq, tags=(), bc_exprs=(), ext=None):
        self.ctx = ctx
        self.dim = dim
        ext = ext if ext is not None else (0.0, 1.0) * dim
        n3 = (C.c_int64 * 3)(*(list(n) + [1] * (3 - dim)))
        e6 = (C.c_double * 6)(*(list(ext) + [0.0] * (6 - len(ext))))
        p3 = (C.c_int32 * 3)(*(list(p) + [1] * (3 - dim)))
        t = (C.c_int32 * max(1, len(tags)))(*tags)
        self._exprs = _cstrs(list(bc_exprs))
        self.h = _vp()
        check(lib().rsd_problem_create_block(ctx.h, C.c_int32(dim), n3, e6, p3, C.c_int32(nq), C.c_int32(len(tags)),
                                             t, self._exprs, C.byref(self.h)))
        self._finish()

    @classmethod
    def from_mesh(cls, ctx, dim, p, nq, xyz, ien, rng, tags=(), bc_exprs=()):
        """A problem on host mesh arrays in the reference's mesh/Data layout:
        xyz [n_nodes*dim], ien [n_elem*npe] (int64), rng [n_elem*2*dim]."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.dim = dim
        xyz = np.ascontiguousarray(xyz, np.float64)
        ien = np.ascontiguousarray(ien, np.int64)
        rng = np.ascontiguousarray(rng, np.int32)
        npe = int(np.prod([q + 1 for q in p]))
        p3 = (C.c_int32 * 3)(*(list(p) + [1] * (3 - dim)))
        t = (C.c_int32 * max(1, len(tags)))(*tags)
        self._exprs = _cstrs(list(bc_exprs))
        self.h = _vp()
        check(lib().rsd_problem_create(ctx.h, C.c_int32(dim), p3, C.c_int32(nq), C.c_int64(xyz.size // dim),
                                       C.c_int64(ien.size // npe), xyz.ctypes.data_as(_d), ien.ctypes.data_as(_i64),
                                       rng.ctypes.data_as(_i32), C.c_int32(len(tags)), t, self._exprs,
                                       C.byref(self.h)))
        self._finish()
        return self

    @classmethod
    def from_pmsh(cls, ctx, path, nq, tags=(), bc_exprs=()):
        """A problem on the mesh of a PMSH file (HeatDispatcher's mesh file)."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.dim = read_pmsh(path)["pdim"]
        t = (C.c_int32 * max(1, len(tags)))(*tags)
        self._exprs = _cstrs(list(bc_exprs))
        self.h = _vp()
        check(lib().rsd_problem_create_pmsh(ctx.h, os.fsencode(path), C.c_int32(nq), C.c_int32(len(tags)), t,
                                            self._exprs, C.byref(self.h)))
        self._finish()
        return self

    def write_pndf(self, path, U, time=0.0):
        """FieldIO::writeBinary: the nodal field of the free vector U with
        the Dirichlet values at `time`."""
        check(lib().rsd_pndf_write(self.ctx.h, self.h, C.c_double(time), dptr(U), os.fsencode(path)))

    def nodal_field(self, U, time=0.0):
        """FieldIO::reconstructNodalField: host values at every node, the
        Dirichlet values at `time` on constrained ones."""
        out = np.zeros(self.n_nodes)
        check(lib().rsd_problem_nodal_field(self.ctx.h, self.h, C.c_double(time), dptr(U), out.ctypes.data_as(_d)))
        return out

    def read_pndf(self, path):
        """FieldIO::readBinary: (nodal values, time); the file's mesh tag must
        be this problem's."""
        nodal = np.zeros(self.n_nodes)
        t = C.c_double()
        check(lib().rsd_pndf_read(self.h, os.fsencode(path), nodal.ctypes.data_as(_d), C.byref(t)))
        return nodal, t.value

    def eval_expr(self, expr, time=0.0, out=None):
        """The free vector of f(t, x, y, z) at the free DOFs' nodes."""
        out = out if out is not None else self.vector()
        check(lib().rsd_problem_eval_expr(self.ctx.h, self.h, expr.encode(), C.c_double(time), dptr(out)))
        return out

    def mesh_tag(self):
        out = C.c_uint64()
        check(lib().rsd_problem_mesh_tag(self.h, C.byref(out)))
        return out.value

    def _finish(self):
        self.ctx._adopt(self)
        s = (C.c_int64 * 8)()
        check(lib().rsd_problem_sizes(self.h, s))
        (self.n_nodes, self.n_elem, self.npe, self.n_free, self.nnz, self.n_colours, self.slots,
         self.max_row) = [int(v) for v in s]

    def close(self):
        """Release the matrices made on this problem, then the problem (the
        library would otherwise orphan them: usable only for their release)."""
        if self.h:
            for m in list(getattr(self, "_matrices", ())):
                m.close()
            lib().rsd_problem_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- getters ------------------------------------------------------------
    def mesh(self):
        xyz = np.zeros(self.n_nodes * self.dim)
        ien = np.zeros(self.n_elem * self.npe, np.int64)
        rng = np.zeros(self.n_elem * 2 * self.dim, np.int32)
        check(lib().rsd_problem_get_mesh(self.h, xyz.ctypes.data_as(_d), ien.ctypes.data_as(_i64),
                                         rng.ctypes.data_as(_i32)))
        return xyz, ien, rng

    def layout(self):
        """(column mode, patterns, int16 offsets fit, slots); mode 0 absolute
        columns, 1 int16 offsets, 2 shared offset patterns."""
        out = (C.c_int64 * 4)()
        check(lib().rsd_problem_layout(self.h, out))
        return tuple(int(v) for v in out)

    def layout_stencil(self):
        """(stencil points, slots) of the FP64 stencil copy the CG reads:
        27 (hex Q1) or 9 (quad Q1) when every row's offsets lie in one such
        stencil, else (0, 0)."""
        out = (C.c_int64 * 2)()
        check(lib().rsd_problem_layout_stencil(self.h, out))
        return tuple(int(v) for v in out)

    def layout_union(self):
        """(1, slots) when the GMRES products read a slice-union copy (each
        64-row slice's union of column offsets in at most 32 positions), else
        (0, 0)."""
        out = (C.c_int64 * 2)()
        check(lib().rsd_problem_layout_union(self.h, out))
        return tuple(int(v) for v in out)

    def set_assembly(self, mode):
        """Matrix assembly scheme: 1 row gather (default), 0 colour-by-colour
        element scatter; returns 1 when the row gather is in use."""
        n = C.c_int32()
        check(lib().rsd_problem_set_assembly(self.h, C.c_int32(mode), C.byref(n)))
        return n.value

    def set_arithmetic(self, mode):
        """Element arithmetic: ARITH_REFERENCE (default, the reference's
        operation order, bitwise), ARITH_CONTRACTED (the Backward Euler
        form pair in fused multiply-adds), ARITH_UNIFORM (on a block mesh,
        every element matrix from element 0's) or ARITH_FACTORED (the pair's
        operand product on Q1 elements without element matrices)."""
        check(lib().rsd_problem_set_arithmetic(self.h, C.c_int32(mode)))

    def padding_f32(self):
        """(lo, hi): the zeros rsd_spmv_f32_padded needs around x."""
        lo, hi = C.c_int64(), C.c_int64()
        check(lib().rsd_problem_padding_f32(self.h, C.byref(lo), C.byref(hi)))
        return lo.value, hi.value

    def layout_f32(self):
        """Layout of the FP32 copy: (column mode, patterns, upper entries per
        row, slots); mode 3 is the symmetric half storage, else the FP64
        layout's mode."""
        out = (C.c_int64 * 4)()
        check(lib().rsd_problem_layout_f32(self.h, out))
        return tuple(int(v) for v in out)

    def dofs(self):
        t2a = np.zeros(self.n_nodes, np.int32)
        a2t = np.zeros(self.n_free, np.int32)
        ct = np.zeros(self.n_nodes, np.int32)
        check(lib().rsd_problem_get_dofs(self.h, t2a.ctypes.data_as(_i32), a2t.ctypes.data_as(_i32),
                                         ct.ctypes.data_as(_i32)))
        return t2a, a2t, ct

    def pattern(self):
        rp = np.zeros(self.n_free + 1, np.int64)
        ci = np.zeros(self.nnz, np.int64)
        check(lib().rsd_problem_get_pattern(self.h, rp.ctypes.data_as(_i64), ci.ctypes.data_as(_i64)))
        return rp, ci

    def elem_map(self):
        m = np.zeros(self.n_elem * self.npe * self.npe, np.int64)
        check(lib().rsd_problem_get_elem_map(self.h, m.ctypes.data_as(_i64)))
        return m

    def colours(self):
        c = np.zeros(self.n_elem, np.int32)
        o = np.zeros(self.n_elem, np.int64)
        check(lib().rsd_problem_get_colours(self.h, c.ctypes.data_as(_i32), o.ctypes.data_as(_i64)))
        return c, o

    # -- model --------------------------------------------------------------
    def set_conductivity(self, k_scalar=1.0, k_expr=None, dk_expr=None, tensor=None, kappa=None):
        tz = None if tensor is None else np.ascontiguousarray(tensor, np.float64)
        kp = None if kappa is None else np.ascontiguousarray(kappa, np.float64)
        check(lib().rsd_problem_set_conductivity(self.h, C.c_double(k_scalar),
                                                 k_expr.encode() if k_expr else None,
                                                 dk_expr.encode() if dk_expr else None,
                                                 None if tz is None else tz.ctypes.data_as(_d),
                                                 None if kp is None else kp.ctypes.data_as(_d)))

    def set_mass(self, rho=1.0, cp=1.0):
        check(lib().rsd_problem_set_mass(self.h, C.c_double(rho), C.c_double(cp)))

    def set_kappa_device(self, d_kappa):
        check(lib().rsd_problem_set_kappa_device(self.h, dptr(d_kappa)))

    # -- assembly -------------------------------------------------------------
    def matrix(self):
        return Matrix(self)

    def vector(self, fill=None):
        v = self.ctx.empty(self.n_free)
        if fill is not None:
            self.ctx.fill(v, fill)
        return v

    def assemble_matrix(self, K, forms, U=None, Udot=None, time=0.0):
        nf, kinds, coefs = _forms(forms)
        check(lib().rsd_assemble_matrix(self.ctx.h, self.h, C.c_double(time), C.c_int32(nf), kinds, coefs,
                                        dptr(U), dptr(Udot), K.h))

    def assemble_vector(self, F, forms, U=None, Udot=None, operand=None, selector=SEL_VALUE, time=0.0):
        nf, kinds, coefs = _forms(forms)
        check(lib().rsd_assemble_vector(self.ctx.h, self.h, C.c_double(time), C.c_int32(nf), kinds, coefs,
                                        dptr(U), dptr(Udot), dptr(operand), C.c_int32(selector), dptr(F)))

    def assemble_load(self, F, src=None, nodal=None, time=0.0):
        check(lib().rsd_assemble_load(self.ctx.h, self.h, C.c_double(time), src.encode() if src else None,
                                      dptr(nodal), dptr(F)))

    def apply_essential(self, F, forms, time=0.0):
        nf, kinds, coefs = _forms(forms)
        check(lib().rsd_apply_essential(self.ctx.h, self.h, C.c_double(time), C.c_int32(nf), kinds, coefs, dptr(F)))

    def apply_natural(self, F, entries, time=0.0):
        """BoundaryApplicator::applyNaturalBCs: F += the flux terms of the
        tagged faces.  entries: list of (tag, flux), flux either a sequence
        of dim expression strings q_sD(t, x, y, z) or a device array of
        nodal flux vectors [n_nodes * dim]."""
        n = len(entries)
        d = self.dim
        tags = (C.c_int32 * max(n, 1))(*[int(t) for t, _ in entries])
        ex = (C.c_char_p * max(n * d, 1))()
        nod = (C.c_void_p * max(n, 1))()
        for k, (_, fl) in enumerate(entries):
            if isinstance(fl, (list, tuple)):
                if len(fl) != d:
                    raise RsdError(f"a flux needs {d} expressions")
                for i, e in enumerate(fl):
                    ex[k * d + i] = e.encode()
            else:
                nod[k] = dptr(fl).value
        check(lib().rsd_apply_natural(self.ctx.h, self.h, C.c_double(time), C.c_int32(n), tags, ex, nod, dptr(F)))

    def assemble_diagonal(self, diag, forms, U=None, time=0.0):
        """Assembler::assembleDiagonal: the diagonal of the forms' matrix."""
        nf, kinds, coefs = _forms(forms)
        check(lib().rsd_assemble_diagonal(self.ctx.h, self.h, C.c_double(time), C.c_int32(nf), kinds, coefs,
                                          dptr(U), dptr(diag)))

    def boundary_reaction(self, tags, U, Udot=None, src=None, time=0.0):
        """QuantityEvaluator::evaluateBoundaryReaction with the heat pairs
        (source on U, -stiffness on U, -mass on Udot when given): returns
        (result, measure) per tag; the monitor's average is result / measure."""
        t = np.ascontiguousarray(tags, np.int32)
        res, meas = np.zeros(len(t)), np.zeros(len(t))
        check(lib().rsd_boundary_reaction(self.ctx.h, self.h, C.c_double(time), src.encode() if src else None,
                                          dptr(U), dptr(Udot), C.c_int32(len(t)), t.ctypes.data_as(_i32),
                                          res.ctypes.data_as(_d), meas.ctypes.data_as(_d)))
        return res, meas

    def boundary_flux(self, tags, U, time=0.0):
        """QuantityEvaluator::evaluateBoundaryRegistry with HeatFluxIntegrand
        (strong flux): returns (integral, measure) per tag; the Average
        reduction is integral / measure."""
        t = np.ascontiguousarray(tags, np.int32)
        res, meas = np.zeros(len(t)), np.zeros(len(t))
        check(lib().rsd_boundary_flux(self.ctx.h, self.h, C.c_double(time), dptr(U), C.c_int32(len(t)),
                                      t.ctypes.data_as(_i32), res.ctypes.data_as(_d), meas.ctypes.data_as(_d)))
        return res, meas

    def fem_apply(self, y, forms, x, U=None, Udot=None, time=0.0):
        """FEMOperator::apply: y = forms applied to the free vector x."""
        self.assemble_vector(y, forms, U=U, Udot=Udot, operand=x, selector=SEL_OPERAND, time=time)

    def cg_fem(self, b, x, forms, tol=1e-10, max_iters=1000, precond=PC_JACOBI, U=None, Udot=None, time=0.0,
               tol_type=TOL_RELATIVE, hist=False):
        """cg::Solver over the matrix-free FEMOperator; returns the Report
        (and the residual history when hist)."""
        rep = Report()
        h = np.zeros(max_iters + 1) if hist else None
        nf, kinds, coefs = _forms(forms)
        check(lib().rsd_cg_solve_fem(self.ctx.h, self.h, C.c_double(time), C.c_int32(nf), kinds, coefs, dptr(U),
                                     dptr(Udot), dptr(b), dptr(x), C.c_double(tol), C.c_int32(tol_type),
                                     C.c_int64(max_iters), C.c_int32(precond), C.byref(rep),
                                     h.ctypes.data_as(_d) if hist else None))
        return (rep, h) if hist else rep

    # -- drivers ----------------------------------------------------------------
    def be_run(self, U, t0, tf, dt, src=None, tol=1e-8, max_iters=1000, precond=PC_JACOBI, use_graph=True,
               max_steps=None, mixed=False):
        max_steps = max_steps or int(round((tf - t0) / dt)) + 2
        iters = np.zeros(max_steps, np.int64)
        steps = C.c_int64()
        timing = np.zeros(2)
        check(lib().rsd_be_run(self.ctx.h, self.h, C.c_double(t0), C.c_double(tf), C.c_double(dt),
                               src.encode() if src else None, C.c_double(tol), C.c_int64(max_iters),
                               C.c_int32(precond), C.c_int32(int(mixed)), C.c_int32(int(use_graph)), dptr(U),
                               C.c_int64(max_steps),
                               C.byref(steps), iters.ctypes.data_as(_i64), timing.ctypes.data_as(_d)))
        self.last_timing_ms = {"assembly": float(timing[0]), "solve": float(timing[1])}
        return steps.value, iters[:steps.value]


    def newton(self, U, src=None, abs_tol=1e-10, rel_tol=1e-10, max_newton=50, lin_solver=1, lin_tol=1e-12,
               lin_max=5000, krylov=30, precond=PC_JACOBI):
        """Steady Newton; lin_solver 0 CG, 1 GMRES(krylov), 2 mixed-precision
        GMRES(krylov), 3 mixed-precision CG (full FP32 copy of J).  Returns (iteration it converged at or < 0, ||R||
        history, inner iterations per Newton step)."""
        its = C.c_int64()
        rh = np.zeros(max_newton)
        li = np.zeros(max_newton, np.int64)
        check(lib().rsd_newton_run(self.ctx.h, self.h, src.encode() if src else None, C.c_double(abs_tol),
                                   C.c_double(rel_tol), C.c_int64(max_newton), C.c_int32(lin_solver),
                                   C.c_double(lin_tol), C.c_int64(lin_max), C.c_int32(krylov), C.c_int32(precond),
                                   dptr(U), C.byref(its), rh.ctypes.data_as(_d), li.ctypes.data_as(_i64)))
        k = its.value
        return k, rh[:max(k, 0) + 1], li[:max(k, 0)]

    def adjoint(self, d_kappa, U0, t0, dt, n_steps, obs, t_obs, tol=1e-13, max_iters=10000, want_grad=True,
                d_grad=None, d_UN=None):
        """Loss and (optionally) dL/dtheta into d_grad; device pointers or DeviceArrays."""
        obs = np.ascontiguousarray(obs, np.int64)
        t_obs = np.ascontiguousarray(t_obs, np.float64)
        loss = C.c_double()
        check(lib().rsd_adjoint_grad_logk(self.ctx.h, self.h, dptr(d_kappa), C.c_double(t0), C.c_double(dt),
                                          C.c_int64(n_steps), C.c_double(tol), C.c_int64(max_iters), dptr(U0),
                                          C.c_int64(obs.size), obs.ctypes.data_as(_i64), t_obs.ctypes.data_as(_d),
                                          C.c_int32(int(want_grad)), C.byref(loss), dptr(d_grad), dptr(d_UN)))
        return loss.value


class Matrix:
    """rsd_matrix: values on the problem's pattern (sliced ELL in HBM)."""

    def __init__(self, pb):
        self.pb = pb
        self.h = _vp()
        check(lib().rsd_matrix_create(pb.h, C.byref(self.h)))
        pb.ctx._adopt(self)
        if not hasattr(pb, "_matrices"):
            pb._matrices = weakref.WeakSet()
        pb._matrices.add(self)

    def close(self):
        if self.h:
            lib().rsd_matrix_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def values(self):
        out = np.zeros(self.pb.nnz)
        check(lib().rsd_matrix_get_values(self.pb.ctx.h, self.h, out.ctypes.data_as(_d)))
        return out

    def set_values(self, v):
        v = np.ascontiguousarray(v, np.float64)
        check(lib().rsd_matrix_set_values(self.pb.ctx.h, self.h, v.ctypes.data_as(_d)))

    def spmv(self, x, y):
        check(lib().rsd_spmv_f64(self.pb.ctx.h, self.h, dptr(x), dptr(y)))

    def to_stencil(self):
        check(lib().rsd_matrix_to_stencil(self.pb.ctx.h, self.h))

    def spmv_stencil(self, x, y):
        check(lib().rsd_spmv_f64_stencil(self.pb.ctx.h, self.h, dptr(x), dptr(y)))

    def inv_diag(self, out):
        check(lib().rsd_jacobi_inv_diag(self.pb.ctx.h, self.h, dptr(out)))

    def gmres(self, b, x, tol=1e-10, max_iters=1000, krylov=30, precond=PC_JACOBI, tol_type=TOL_RELATIVE):
        rep = Report()
        check(lib().rsd_gmres_solve_f64(self.pb.ctx.h, self.h, dptr(b), dptr(x), C.c_double(tol), C.c_int32(tol_type),
                                        C.c_int64(max_iters), C.c_int32(krylov), C.c_int32(precond), C.byref(rep)))
        return rep

    def gmres_mixed(self, b, x, tol=1e-10, max_iters=1000, krylov=30, precond=PC_JACOBI, inner_floor=0.0):
        """FP32 inner GMRES(krylov) with FP64 refinement; returns (report,
        outer cycles)."""
        rep = Report()
        cyc = C.c_int64()
        check(lib().rsd_gmres_solve_mixed(self.pb.ctx.h, self.h, dptr(b), dptr(x), C.c_double(tol),
                                          C.c_int64(max_iters), C.c_int32(krylov), C.c_int32(precond),
                                          C.c_double(inner_floor), C.byref(rep), C.byref(cyc)))
        return rep, cyc.value

    def to_f32(self):
        check(lib().rsd_matrix_to_f32(self.pb.ctx.h, self.h))

    def spmv_f32(self, x, y):
        check(lib().rsd_spmv_f32(self.pb.ctx.h, self.h, dptr(x), dptr(y)))

    def spmv_f32_padded(self, xp, y):
        """y = A x by the mixed CG's kernel; xp is a float32 DeviceArray of
        lo + n + hi entries (Problem.padding_f32) holding x at [lo, lo + n)
        and zeros around it."""
        lo, hi = self.pb.padding_f32()
        assert xp.n == lo + self.pb.n_free + hi and xp.dtype == np.float32
        check(lib().rsd_spmv_f32_padded(self.pb.ctx.h, self.h, _vp(xp.ptr + 4 * lo), dptr(y)))

    def cg_mixed(self, b, x, tol=1e-8, max_iters=1000, precond=PC_JACOBI, inner_floor=0.0):
        rep = Report()
        cyc = C.c_int64()
        check(lib().rsd_cg_solve_mixed(self.pb.ctx.h, self.h, dptr(b), dptr(x), C.c_double(tol), C.c_int64(max_iters),
                                       C.c_int32(precond), C.c_double(inner_floor), C.byref(rep), C.byref(cyc)))
        return rep, cyc.value

    def cg(self, b, x, tol=1e-10, max_iters=1000, precond=PC_JACOBI, hist=False, tol_type=TOL_RELATIVE):
        rep = Report()
        h = np.zeros(max_iters + 1) if hist else None
        check(lib().rsd_cg_solve_f64(self.pb.ctx.h, self.h, dptr(b), dptr(x), C.c_double(tol), C.c_int32(tol_type),
                                     C.c_int64(max_iters), C.c_int32(precond), C.byref(rep),
                                     None if h is None else h.ctypes.data_as(_d)))
        return (rep, h) if hist else rep


class Batch:
    """rsd_batch: independent systems solved by one set of launches."""

    def __init__(self, ctx, mats):
        self.ctx, self.mats = ctx, list(mats)
        arr = (_vp * len(self.mats))(*[m.h.value for m in self.mats])
        self.h = _vp()
        check(lib().rsd_batch_create(ctx.h, C.c_int32(len(self.mats)), arr, C.byref(self.h)))
        ctx._adopt(self)
        tot = C.c_int64()
        self.offsets = np.zeros(len(self.mats) + 1, np.int64)
        check(lib().rsd_batch_layout(self.h, C.byref(tot), self.offsets.ctypes.data_as(_i64)))
        self.total_rows = tot.value

    def close(self):
        if self.h:
            lib().rsd_batch_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pack(self, vecs):
        out = np.zeros(self.total_rows)
        for i, v in enumerate(vecs):
            out[self.offsets[i]:self.offsets[i] + v.size] = v
        return out

    def unpack(self, packed):
        return [packed[self.offsets[i]:self.offsets[i] + m.pb.n_free].copy() for i, m in enumerate(self.mats)]

    @property
    def schedule(self):
        """(chunk, unroll, order, xcd): the batch's launch schedule."""
        s = (C.c_int32 * 4)()
        check(lib().rsd_batch_get_schedule(self.h, s))
        return tuple(int(v) for v in s)

    @schedule.setter
    def schedule(self, sched):
        check(lib().rsd_batch_set_schedule(self.h, (C.c_int32 * 4)(*[int(v) for v in sched])))

    def autotune(self, precond=PC_JACOBI, cache=None, iters=64):
        """Time every schedule on this batch and keep the fastest (a cached
        one when `cache`, a JSON file, holds this batch's key); returns
        (schedule, us per iteration, from_cache)."""
        s = (C.c_int32 * 4)()
        us = C.c_double()
        hit = C.c_int32()
        check(lib().rsd_batch_autotune(self.ctx.h, self.h, C.c_int32(precond), cache.encode() if cache else None,
                                       C.c_int64(iters), s, C.byref(us), C.byref(hit)))
        return tuple(int(v) for v in s), us.value, bool(hit.value)

    def cg(self, b, x, tol=1e-8, max_iters=1000, precond=PC_JACOBI, use_graph=True):
        reps = (Report * len(self.mats))()
        check(lib().rsd_cg_solve_batched(self.ctx.h, self.h, dptr(b), dptr(x), C.c_double(tol), C.c_int64(max_iters),
                                         C.c_int32(precond), C.c_int32(int(use_graph)), reps))
        return list(reps)

    def cg_mixed(self, b, x, tol=1e-8, max_iters=1000, precond=PC_JACOBI, inner_floor=0.0):
        """Every problem's mixed-precision solve (FP32 inner CG, FP64 outer
        residual) in one set of launches; returns (reports, outer cycles)."""
        reps = (Report * len(self.mats))()
        cyc = np.zeros(len(self.mats), np.int64)
        check(lib().rsd_cg_solve_batched_mixed(self.ctx.h, self.h, dptr(b), dptr(x), C.c_double(tol),
                                               C.c_int64(max_iters), C.c_int32(precond), C.c_double(inner_floor),
                                               reps, cyc.ctypes.data_as(_i64)))
        return list(reps), [int(c) for c in cyc]


def philox_normal(seed, stream, n):
    out = np.zeros(n)
    check(lib().rsd_philox_normal(C.c_uint64(seed), C.c_uint32(stream), C.c_int64(n), out.ctypes.data_as(_d)))
    return out


def philox_uniform(seed, stream, n):
    out = np.zeros(n)
    check(lib().rsd_philox_uniform(C.c_uint64(seed), C.c_uint32(stream), C.c_int64(n), out.ctypes.data_as(_d)))
    return out


def exported_symbols():
    """Names declared in include/residuum_hip.h (for the load test)."""
    import re
    hdr = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include", "residuum_hip.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(rsd_\w+)\s*\(", text, re.M)))
