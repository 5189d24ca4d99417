"""Oracle: CPU restatement of OpFlow's hot path -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker. Two parts:
  * liboracle.so (oracle.c): ranges, updatePadding, FTCS with OpFlow's
    FieldAssigner sequence, reductions, CSR assembly and the CPU PCG solver;
  * expr.py: numpy evaluator of OpFlow expression trees (lazy evalAt
    recursion of Expr/Expression.hpp:51-65 with each operator's eval).
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
        _lib.orc_ftcs.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_int]
        _lib.orc_update_padding.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_desc_finalize.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_reduce.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        _lib.orc_mesh_uniform.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_void_p]
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


from .expr import evaluate, assign  # noqa: E402,F401
