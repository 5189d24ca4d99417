"""Implicit-solve front-end: StructSolverParams, Solve, makeEqnSolveHandler
(Equation/UnifiedSolve.hpp:36-67, Equation/HYPREEqnSolveHandler.hpp:44-48),
plus direct multigrid and serving handles over the C ABI."""
import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib as L


@dataclass
class StructSolverParams:
    """StructSolverParamsBase (Solvers/Struct/StructSolver.hpp:39-51) + engine knobs."""
    type: int = 6  # StructSolverType::GMRES (UnifiedSolve.hpp:36 default)
    tol: Optional[float] = None
    maxIter: Optional[int] = None
    staticMat: bool = False
    pinValue: bool = False
    precision: str = "fp64"  # "fp64" | "mixed"
    numPreRelax: Optional[int] = None
    numPostRelax: Optional[int] = None
    verbose: bool = False
    # HYPRE knobs: honoured where they map onto the engine, refused where they would change
    # the stopping test or the smoother kind, accepted without effect where they only select
    # PFMG internals (rapType, skipRelax) -- as the C++ front-end does
    useZeroGuess: Optional[bool] = None
    relaxType: Optional[int] = None
    rapType: Optional[int] = None
    skipRelax: Optional[int] = None
    twoNorm: Optional[int] = None
    relChange: Optional[int] = None
    absTol: Optional[float] = None


def _check_params(q, which):
    """Refuse (OPF_ERR_UNSUPPORTED) the HYPRE settings the engine cannot honour."""
    if q is None:
        return
    msg = None
    if q.absTol is not None and q.absTol > 0:
        msg = "absTol > 0 is not supported (the engine stops on tol, the relative residual ||b-Ax||/||b||)"
    elif q.twoNorm is not None and q.twoNorm == 0:
        msg = "twoNorm = 0 is not supported (the stopping test uses the 2-norm)"
    elif q.relChange is not None and q.relChange != 0:
        msg = "relChange is not supported"
    elif q.relaxType is not None and q.relaxType not in (0, 1):
        msg = "relaxType >= 2 (Gauss-Seidel) is not supported: the smoother is weighted Jacobi"
    if msg:
        raise L.OpfError(L.OPF_ERR_UNSUPPORTED, f"{which}: {msg}")


@dataclass
class EqnSolveState:  # Equation/EqnSolveHandler.hpp:17-25
    niter: int = 0
    relerr: float = 0.0
    abserr: float = 0.0
    nrefine: int = 0
    converged: bool = False


def _params(params, precParams):
    _check_params(params, "solver params")
    _check_params(precParams, "preconditioner params")
    p = L.SolveParams()
    p.solver = int(params.type if params else 6)
    p.precond = int(precParams.type if precParams else 0)
    p.tol = float(params.tol) if params and params.tol is not None else 0.0
    p.max_iter = int(params.maxIter) if params and params.maxIter is not None else 0
    p.static_mat = int(bool(params and params.staticMat))
    p.pin_value = int(bool(params and params.pinValue))
    p.precision = 1 if params and params.precision == "mixed" else 0
    src = precParams or params
    p.pre_smooth = int(src.numPreRelax) if src and src.numPreRelax else 0
    p.post_smooth = int(src.numPostRelax) if src and src.numPostRelax else 0
    p.verbose = int(bool(params and params.verbose))
    return p


class PrecondStructSolver:
    """PrecondStructSolver<type, pType>(params, precParams) (StructSolverPrecond.hpp:19-71)."""

    def __init__(self, params=None, precParams=None):
        self.params = params or StructSolverParams()
        self.precParams = precParams


class _Handler:
    """EqnSolveHandler whose solve() runs on the device (EqnSolveHandler.hpp:28-33)."""

    def __init__(self, getter, target, solver):
        from . import Equation, _node_array  # noqa: F401
        from . import Expr
        eqn = getter(target)
        if not hasattr(eqn, "lhs"):
            raise TypeError("equation getter must return `lhs == rhs`")
        # one combined post-order list holding both sides
        both = Expr(3, eqn.lhs, eqn.rhs)  # lhs - rhs, root's children are the two sides
        nodes, fields, root = both.flatten(unknown=target)
        lhs_root, rhs_root = nodes[root][1], nodes[root][2]
        nodes = nodes[:-1]
        self.fields = [target] + [f for f in fields if f is not target]
        slot = {id(f): i for i, f in enumerate(self.fields)}
        remap = []
        for (k, a, b, d, v) in nodes:
            if k == 0:
                a = slot[id(fields[a])]
            remap.append((k, a, b, d, v))
        from . import _node_array as na
        self._nodes = na(remap)
        fh = (C.c_void_p * len(self.fields))(*[f.h.value for f in self.fields])
        self.params = _params(solver.params, solver.precParams)
        self.target = target
        self.zero_guess = bool(solver.params and solver.params.useZeroGuess)
        h = C.c_void_p()
        lib = target.lib
        L.check(lib.opf_eqn_create(self._nodes, len(remap), lhs_root, rhs_root, fh, len(self.fields), target.h,
                                   C.byref(self.params), C.byref(h)))
        self.h, self.lib = h, lib

    def solve(self):
        st = L.SolveState()
        if self.zero_guess:
            self.target.assign(0.0)
        L.check(self.lib.opf_eqn_solve(self.h, C.byref(st)))
        return EqnSolveState(st.niter, st.relerr, st.abserr, st.nrefine, bool(st.converged))

    def csr(self):
        """(ptr, col, val, rhs) of the system as CSRMatrixGenerator builds it (pin last when pinValue)."""
        nr, nz = C.c_longlong(), C.c_longlong()
        L.check(self.lib.opf_eqn_csr(self.h, C.byref(nr), C.byref(nz), None, None, None, None))
        ptr = np.zeros(nr.value + 1, dtype=np.int64)
        col = np.zeros(nz.value, dtype=np.int64)
        val = np.zeros(nz.value)
        rhs = np.zeros(nr.value)
        L.check(self.lib.opf_eqn_csr(self.h, C.byref(nr), C.byref(nz), ptr.ctypes.data, col.ctypes.data,
                                     val.ctypes.data, rhs.ctypes.data))
        return ptr, col, val, rhs

    def close(self):
        if self.h:
            self.lib.opf_eqn_destroy(self.h)
            self.h = None


def makeEqnSolveHandler(getter, target, solver):
    return _Handler(getter, target, solver)


def Solve(getter, target, params=None, precParams=None):
    """Solve(func, target, params, precParams) (UnifiedSolve.hpp:36-43)."""
    h = _Handler(getter, target, PrecondStructSolver(params, precParams))
    try:
        return h.solve()
    finally:
        h.close()


@dataclass
class IJSolverParams(StructSolverParams):
    """IJSolverParams (Solvers/IJ/IJSolver.hpp:24-71) as SolveEqns reads them: params[0]'s
    tol / maxIter for the whole system (amgcl defaults 1e-8 / 100 when unset), pinValue
    per equation."""


class ColoredMDRangeMapper:
    """DS::ColoredMDRangeMapper (DataStructures/Index/LinearMapper/ColoredMDRangeMapper.hpp:20-41):
    block i ranks range i's cells x-fastest after the cells of blocks 0..i-1."""

    def __init__(self, *ranges):
        self.ranges = list(ranges)
        self.offset = [0]
        for r in self.ranges:
            n = 1
            for lo, hi in r:
                n *= hi - lo
            self.offset.append(self.offset[-1] + n)

    def __call__(self, idx, color):
        r = self.ranges[color]
        rank, stride = 0, 1
        for (lo, hi), i in zip(r, idx):
            rank += (i - lo) * stride
            stride *= hi - lo
        return self.offset[color] + rank


class _EqnHolder:
    """EqnHolder / StencilHolder (Equation/EquationHolder.hpp:31-99): each getter is called
    with every target's StencilField placeholder; the lowered trees go to the device."""

    def __init__(self, getters, targets):
        from . import Expr, StencilUnknown, _node_array
        if len(getters) != len(targets) or not getters:
            raise ValueError("an equation set needs one getter per target")
        self.targets = list(targets)
        unk = [StencilUnknown(i, t) for i, t in enumerate(self.targets)]
        n = len(getters)
        self.descs = (L.EqnDesc * n)()
        self._keep = []
        for i, g in enumerate(getters):
            eqn = g(*unk)
            if not hasattr(eqn, "lhs"):
                raise TypeError("equation getter must return `lhs == rhs`")
            both = Expr(3, eqn.lhs, eqn.rhs)
            nodes, fields, root = both.flatten()
            lhs_root, rhs_root = nodes[root][1], nodes[root][2]
            nodes = nodes[:-1]
            arr = _node_array(nodes)
            fh = (C.c_void_p * max(1, len(fields)))(*[f.h.value for f in fields])
            self._keep += [arr, fh, fields]
            d = self.descs[i]
            d.nodes = C.cast(arr, C.POINTER(L.ExprNode))
            d.nnodes = len(nodes)
            d.lhs_root, d.rhs_root = lhs_root, rhs_root
            d.fields = C.cast(fh, C.POINTER(C.c_void_p))
            d.nfields = len(fields)
        self.th = (C.c_void_p * n)(*[t.h.value for t in self.targets])
        self.lib = self.targets[0].lib

    @property
    def size(self):
        return len(self.targets)


def makeEqnHolder(getters, targets):
    return _EqnHolder(list(getters), list(targets))


def makeStencilHolder(eqn_holder):
    return eqn_holder


@dataclass
class CSRMatrix:  # DS::CSRMatrix (DataStructures/Matrix/CSRMatrix.hpp)
    row: np.ndarray
    col: np.ndarray
    val: np.ndarray
    rhs: np.ndarray


class CSRMatrixGenerator:
    """CSRMatrixGenerator::generate (Equation/CSRMatrixGenerator.hpp:26-41, 143-193) of a set:
    pads evaluated on the device, rows in ColoredMDRangeMapper order, pin_flags[i] makes
    block i's last row x = 0."""

    @staticmethod
    def generate(st_holder, mapper=None, pin_flags=None):
        h = st_holder
        n = h.size
        if mapper is not None and len(mapper.ranges) != n:
            raise ValueError("the mapper needs one range per target")
        pins = (C.c_int * n)(*[int(bool(p)) for p in (pin_flags or [False] * n)])
        nr, nz = C.c_longlong(), C.c_longlong()
        L.check(h.lib.opf_eqns_csr(n, h.descs, h.th, pins, C.byref(nr), C.byref(nz), None, None, None, None))
        ptr = np.zeros(nr.value + 1, dtype=np.int64)
        col = np.zeros(nz.value, dtype=np.int64)
        val = np.zeros(nz.value)
        rhs = np.zeros(nr.value)
        L.check(h.lib.opf_eqns_csr(n, h.descs, h.th, pins, C.byref(nr), C.byref(nz), ptr.ctypes.data,
                                   col.ctypes.data, val.ctypes.data, rhs.ctypes.data))
        return CSRMatrix(ptr, col, val, rhs)


def SolveEqns(*args):
    """SolveEqns<Solver>(getters..., targets..., mapper, params) (Equation/UnifiedSolve.hpp:69-100):
    solves the set for every target from x = 0 and writes them back (padding updated)."""
    if len(args) < 4 or len(args) % 2:
        raise TypeError("SolveEqns(getters..., targets..., mapper, params)")
    n = (len(args) - 2) // 2
    getters, targets, mapper, params = args[:n], args[n:2 * n], args[2 * n], args[2 * n + 1]
    if mapper is not None and len(mapper.ranges) != n:
        raise ValueError("the mapper needs one range per target")
    params = list(params) if params is not None else [IJSolverParams() for _ in range(n)]
    if len(params) != n:
        raise ValueError("SolveEqns needs one parameter set per equation")
    h = _EqnHolder(list(getters), list(targets))
    pa = (L.SolveParams * n)(*[_params(p, None) for p in params])
    st = L.SolveState()
    L.check(h.lib.opf_eqns_solve(n, h.descs, h.th, pa, C.byref(st)))
    return EqnSolveState(st.niter, st.relerr, st.abserr, st.nrefine, bool(st.converged))


def gen_bumps_grid(desc, seed, nbumps, noise=0.0):
    """RHS of seeded bumps (+ noise) on the unit-cube cell centres of a Poisson descriptor."""
    n = 1
    for q in range(desc.dim):
        n *= desc.n[q]
    out = np.zeros(n)
    L.check(L.load().opf_gen_bumps_grid(C.byref(desc), C.c_uint64(seed), nbumps, noise, out.ctypes.data_as(L._DP)))
    return out


def gen_bumps_grid_dev(ctx, desc, seed, nbumps, noise, buf):
    """gen_bumps_grid's field written on the device into DeviceBuffer `buf` (context stream)."""
    L.check(ctx.lib.opf_gen_bumps_grid_dev(ctx.h, C.byref(desc), C.c_uint64(seed), nbumps, noise, buf.ptr))


def poisson_desc(dim, n, h, bc, sigma=0.0):
    d = L.PoissonDesc()
    d.dim = dim
    for i in range(3):
        d.n[i] = n[i] if i < dim else 1
        d.h[i] = h[i] if i < dim else 1.0
        for s in range(2):
            d.bc[i][s] = int(bc[i][s]) if i < dim else 0
    d.sigma = sigma
    return d


class DeviceBuffer:
    """Raw device array of float64 owned by the library (opf_buf_*)."""

    def __init__(self, ctx, n):
        self.ctx, self.n, self.lib = ctx, n, ctx.lib
        p = C.c_void_p()
        L.check(self.lib.opf_buf_create(ctx.h, 8 * n, C.byref(p)))
        self.ptr = p

    def upload(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        assert a.size == self.n
        L.check(self.lib.opf_buf_upload(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes))

    def numpy(self):
        out = np.empty(self.n)
        L.check(self.lib.opf_buf_download(self.ctx.h, self.ptr, out.ctypes.data, out.nbytes))
        return out

    def zero(self):
        L.check(self.lib.opf_buf_zero(self.ctx.h, self.ptr, self.n * 8))

    def close(self):
        if self.ptr:
            self.lib.opf_buf_destroy(self.ctx.h, self.ptr)
            self.ptr = None


class PoissonGMG:
    """Direct matrix-free GMG-PCG on device vectors (opf_gmg_*)."""

    def __init__(self, ctx, desc, precision="fp64"):
        self.lib, self.ctx, self.desc = ctx.lib, ctx, desc
        h = C.c_void_p()
        L.check(self.lib.opf_gmg_create(ctx.h, C.byref(desc), 1 if precision == "mixed" else 0, C.byref(h)))
        self.h = h

    def solve(self, b_ptr, x_ptr, tol=1e-10, max_iter=200, pre=0, post=0):
        p = L.SolveParams()
        p.solver, p.precond, p.tol, p.max_iter, p.pre_smooth, p.post_smooth = 5, 3, tol, max_iter, pre, post
        st = L.SolveState()
        L.check(self.lib.opf_gmg_solve(self.h, b_ptr, x_ptr, C.byref(p), C.byref(st)))
        return EqnSolveState(st.niter, st.relerr, st.abserr, st.nrefine, bool(st.converged))

    def apply(self, x_ptr, y_ptr):
        L.check(self.lib.opf_gmg_apply(self.h, x_ptr, y_ptr))

    def close(self):
        if self.h:
            self.lib.opf_gmg_destroy(self.h)
            self.h = None


class Server:
    """Continuous-batching Poisson/heat solve server (opf_srv_*)."""

    def __init__(self, ctx, mem_budget, max_inflight=8, precision="mixed", batched=True, tol=1e-8, max_iter=100):
        self.lib = ctx.lib
        cfg = L.SrvConfig()
        cfg.mem_budget, cfg.max_inflight = int(mem_budget), max_inflight
        cfg.precision, cfg.batched, cfg.tol, cfg.max_iter = (1 if precision == "mixed" else 0), int(batched), tol, \
            max_iter
        h = C.c_void_p()
        L.check(self.lib.opf_srv_create(ctx.h, C.byref(cfg), C.byref(h)))
        self.h = h
        self._outs = []  # host buffers the server writes into

    def submit(self, desc, seed, nbumps, noise=0.0, rhs=None, out=None, t_submit=-1.0):
        """Queue a solve; `out` (float64 array of prod(n)) receives the solution when it finishes."""
        r = L.SrvRequest()
        r.desc, r.seed, r.nbumps, r.noise, r.t_submit = desc, seed, nbumps, noise, t_submit
        keep = []
        if rhs is not None:
            a = np.ascontiguousarray(rhs, dtype=np.float64)
            keep.append(a)
            r.rhs = a.ctypes.data_as(L._DP)
        if out is not None:
            assert out.dtype == np.float64 and out.flags["C_CONTIGUOUS"]
            self._outs.append(out)
            r.out_x = out.ctypes.data_as(L._DP)
        i = C.c_uint64()
        L.check(self.lib.opf_srv_submit(self.h, C.byref(r), C.byref(i)))
        return i.value

    def step(self):
        n = C.c_int()
        L.check(self.lib.opf_srv_step(self.h, C.byref(n)))
        return n.value

    def poll(self, maxn=256):
        arr = (L.SrvResult * maxn)()
        n = C.c_int()
        L.check(self.lib.opf_srv_poll(self.h, arr, maxn, C.byref(n)))
        return [arr[i] for i in range(n.value)]

    def pending(self):
        q, f = C.c_int(), C.c_int()
        L.check(self.lib.opf_srv_pending(self.h, C.byref(q), C.byref(f)))
        return q.value, f.value

    def clock(self):
        return self.lib.opf_srv_clock(self.h)

    def close(self):
        if self.h:
            self.lib.opf_srv_destroy(self.h)
            self.h = None
