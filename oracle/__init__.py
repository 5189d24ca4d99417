"""Oracle: CPU restatement of OpFlow's hot path -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker. Three parts:
  * liboracle.so (oracle.c): ranges, updatePadding, FTCS with OpFlow's
    FieldAssigner sequence, reductions, CSR assembly and the CPU PCG solver;
  * expr.py: numpy evaluator of OpFlow expression trees (lazy evalAt
    recursion of Expr/Expression.hpp:51-65 with each operator's eval), and
    linear_system, the CSRMatrixGenerator system of an equation set taken by
    linearity from that evaluator;
  * cart.py: the byte layout of the raw binary field streams.
Each function's pin (golden arrays, closed forms, documented structure) is
named in the test that checks it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so missing: run __graft_entry__.build()")
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_reduce.restype = C.c_double
        _lib.orc_num_threads.restype = C.c_int
        _lib.orc_set_threads.argtypes = [C.c_int]
        _lib.orc_ftcs.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_int]
        _lib.orc_update_padding.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_desc_finalize.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_reduce.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        _lib.orc_mesh_uniform.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_void_p]
        _lib.orc_csr_poisson.restype = C.c_long
        _lib.orc_csr_poisson.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p]
        _lib.orc_poisson_apply.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                                           C.c_void_p]
        _lib.orc_pcg_mg.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                                    C.c_double, C.c_int, C.c_void_p]
    return _lib


class OrcMesh(C.Structure):
    _fields_ = [("x", C.c_void_p * 3), ("dx", C.c_void_p * 3), ("org", C.c_int * 3)]


class Mesh:
    """Host mesh arrays (uniform MeshBuilder restatement) for the oracle."""

    def __init__(self, dims, ranges=None, pad=5):
        self.dim = len(dims)
        self.dims = list(dims)
        self.pad = pad
        ranges = ranges or [(0.0, 1.0)] * self.dim
        self.xs, self.dxs = [], []
        for d in range(self.dim):
            x = np.zeros(dims[d] + 2 * pad)
            dx = np.zeros(dims[d] + 2 * pad - 1)
            lib().orc_mesh_uniform(dims[d], ranges[d][0], ranges[d][1], pad, x.ctypes.data, dx.ctypes.data)
            self.xs.append(x)
            self.dxs.append(dx)
        self.s = OrcMesh()
        for d in range(self.dim):
            self.s.x[d] = self.xs[d].ctypes.data
            self.s.dx[d] = self.dxs[d].ctypes.data
        for d in range(3):
            self.s.org[d] = -pad if d < self.dim else 0

    def x(self, d, i):
        return self.xs[d][np.asarray(i) + self.pad]

    def dx(self, d, i):
        return self.dxs[d][np.asarray(i) + self.pad]


def finalize(desc_struct, mesh):
    """orc_desc_finalize on a copy of an opflow_amd FieldDesc-compatible struct."""
    d = type(desc_struct).from_buffer_copy(desc_struct)
    start = (C.c_int * 3)(0, 0, 0)
    nn = (C.c_int * 3)(*(mesh.dims + [1] * (3 - mesh.dim)))
    rc = lib().orc_desc_finalize(C.byref(d), start, nn)
    assert rc == 0
    return d


def update_padding(desc, mesh, arr):
    a = np.ascontiguousarray(arr, dtype=np.float64)
    lib().orc_update_padding(C.byref(desc), C.byref(mesh.s), a.ctypes.data)
    return a


def ftcs(desc, mesh, arr, coef, nsteps, nthreads=0):
    a = np.ascontiguousarray(arr, dtype=np.float64).copy()
    lib().orc_ftcs(C.byref(desc), C.byref(mesh.s), a.ctypes.data, coef, nsteps, nthreads)
    return a


def reduce(desc, arr, red, rs, re_):
    a = np.ascontiguousarray(arr, dtype=np.float64)
    r = {"sum": 0, "max": 1, "min": 2, "sumsq": 3, "maxabs": 4}[red]
    rs3 = (C.c_int * 3)(*(list(rs) + [0] * (3 - len(rs))))
    re3 = (C.c_int * 3)(*(list(re_) + [1] * (3 - len(re_))))
    return lib().orc_reduce(C.byref(desc), a.ctypes.data, r, rs3, re3)


def num_threads():
    return lib().orc_num_threads()


def set_threads(n):
    """OpenMP threads of later oracle calls; n <= 0 restores the process's initial count."""
    lib().orc_set_threads(int(n))


def csr_poisson(desc, mesh, c=1.0, pin_last=False):
    """CSR (ptr, col, val, rhs) of `c == sum_d d2<D2SecondOrderCentered>(e)` (CSRMatrixGenerator restated)."""
    ncell = 1
    for d in range(desc.dim):
        ncell *= desc.access_end[d] - desc.access_start[d]
    ptr = np.zeros(ncell + 1, dtype=np.int64)
    col = np.zeros(ncell * 32, dtype=np.int64)
    val = np.zeros(ncell * 32)
    rhs = np.zeros(ncell)
    nnz = lib().orc_csr_poisson(C.byref(desc), C.byref(mesh.s), c, int(pin_last), ptr.ctypes.data, col.ctypes.data,
                                val.ctypes.data, rhs.ctypes.data)
    return ptr, col[:nnz], val[:nnz], rhs


def _pois_args(n, h, bc):
    dim = len(n)
    n3 = (C.c_int * 3)(*(list(n) + [1] * (3 - dim)))
    h3 = (C.c_double * 3)(*(list(h) + [1.0] * (3 - dim)))
    bc3 = ((C.c_int * 2) * 3)()
    for d in range(dim):
        bc3[d][0], bc3[d][1] = bc[d]
    return dim, n3, h3, bc3


def poisson_apply(n, h, bc, x, sigma=0.0):
    dim, n3, h3, bc3 = _pois_args(n, h, bc)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    lib().orc_poisson_apply(dim, n3, h3, bc3, sigma, x.ctypes.data, y.ctypes.data)
    return y


def pcg_mg(n, h, bc, b, x0=None, sigma=0.0, tol=1e-10, maxit=200):
    dim, n3, h3, bc3 = _pois_args(n, h, bc)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).copy()
    rel = C.c_double()
    it = lib().orc_pcg_mg(dim, n3, h3, bc3, sigma, b.ctypes.data, x.ctypes.data, tol, maxit, C.byref(rel))
    return x, it, rel.value


REF_PATH = os.path.join(_HERE, "_ref", "libamgcl_ref.so")


def amgcl_solve(ptr, col, val, rhs, tol=1e-12, maxiter=500, kind=0):
    """Reference solve with AMGCL (oracle/_ref, built from the reference tree's headers)."""
    if not os.path.exists(REF_PATH):
        return None
    ref = C.CDLL(REF_PATH)
    ref.amgcl_ref_solve.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_double, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    n = len(rhs)
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int64)
    val = np.ascontiguousarray(val, dtype=np.float64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    x = np.zeros(n)
    it, res = C.c_int(), C.c_double()
    ref.amgcl_ref_solve(n, ptr.ctypes.data, col.ctypes.data, val.ctypes.data, rhs.ctypes.data, x.ctypes.data, tol,
                        maxiter, kind, C.byref(it), C.byref(res))
    return x, it.value, res.value


from .expr import evaluate, assign  # noqa: E402,F401
