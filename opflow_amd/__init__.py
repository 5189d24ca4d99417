"""opflow_amd -- host-side mirror of OpFlow's explicit/implicit front-end over
the MI355X engine (libopflow_hip.so, C ABI in include/opflow_hip.h).

Names follow OpFlow (src/Core/...): MeshBuilder/CartesianMesh, ExprBuilder/
CartesianField with setBC/setExt/setLoc/setPadding, the operator kernels
D2SecondOrderCentered, D1FirstOrderCentered, D1FirstOrderBiasedUpwind/
Downwind, D1WENO53Upwind/Downwind, D1Linear interpolation, `field.assign(expr)`
for OpFlow's `field = expr`, rangeReduce, Solve/makeEqnSolveHandler and
StructSolverParams. Every compute call goes to the HIP library; there is no
CPU path in this package.
"""
import builtins
import ctypes as C
import enum

import numpy as np

from . import _lib as L

__all__ = [
    "Context", "LocOnMesh", "BCType", "DimPos", "BasicArithOp", "StructSolverType", "MeshBuilder",
    "CartesianMesh", "ExprBuilder", "CartesianField", "Expr", "D2SecondOrderCentered", "D1FirstOrderCentered",
    "D1FirstOrderBiasedUpwind", "D1FirstOrderBiasedDownwind", "D1WENO53Upwind", "D1WENO53Downwind",
    "d2x", "d2y", "d2z", "dx", "dy", "dz", "d1", "d2", "d1IntpCenterToCorner", "d1IntpCornerToCenter",
    "rangeReduce", "StructSolverParams", "EqnSolveState", "Solve", "makeEqnSolveHandler", "PrecondStructSolver",
    "PoissonGMG", "Server", "SolveEqns", "makeEqnHolder", "makeStencilHolder", "CSRMatrixGenerator",
    "ColoredMDRangeMapper", "IJSolverParams", "StencilUnknown", "TimeStamp", "RawBinaryOStream",
    "RawBinaryIStream", "Utils",
]


class LocOnMesh(enum.IntEnum):  # src/Core/Constants.hpp:50
    Corner = 0
    Center = 1


class BCType(enum.IntEnum):  # src/Core/BC/BCBase.hpp:24
    Undefined = 0
    Dirc = 1
    Neum = 2
    Periodic = 3
    Internal = 4
    Symm = 5
    ASymm = 6


class DimPos(enum.IntEnum):  # src/Core/Constants.hpp:48
    start = 0
    end = 1


class BasicArithOp(enum.IntEnum):  # src/Core/Constants.hpp:53
    Eq = 0
    Add = 1
    Minus = 2
    Mul = 3
    Div = 4


class StructSolverType(enum.IntEnum):  # src/Core/Solvers/Struct/StructSolver.hpp:23-34
    None_ = 0
    Jacobi = 1
    SMG = 2
    PFMG = 3
    CYCRED = 4
    PCG = 5
    GMRES = 6
    FGMRES = 7
    LGMRES = 8
    BICGSTAB = 9


# ---------------------------------------------------------------- context
class Context:
    """One device + one HIP stream (replaces the global ParallelPlan/TBB arena)."""

    def __init__(self, device=0):
        self.lib = L.load()
        h = C.c_void_p()
        L.check(self.lib.opf_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def sync(self):
        L.check(self.lib.opf_ctx_sync(self.h))

    def set_jit(self, mode):
        """Engine of expression sweeps: 0 interpreter, 1 compiled kernels for >= 2^18 cells
        (default), 2 compiled kernels always, 3 / 4 compiled kernels always in the LDS-staged /
        direct form (same bits in every mode)."""
        L.check(self.lib.opf_ctx_set_jit(self.h, int(mode)))

    def cache_stats(self):
        """Counters of the sweep caches (opf_cache_stats): hipRTC compiles and compiled kernels
        loaded in the process, kernels loaded now, programs this context uploaded, resident now."""
        a = (C.c_longlong * 5)()
        L.check(self.lib.opf_cache_stats(self.h, a))
        keys = ("jit_compiles", "jit_loaded", "jit_resident", "programs_uploaded", "programs_resident")
        return dict(zip(keys, list(a)))

    @property
    def stream(self):
        return self.lib.opf_ctx_stream(self.h)

    def set_stream(self, stream_ptr):
        """Enqueue this context's work on a caller-owned HIP stream (an int / c_void_p
        hipStream_t, e.g. torch.cuda.current_stream().cuda_stream); None: own stream."""
        L.check(self.lib.opf_ctx_set_stream(self.h, C.c_void_p(stream_ptr) if stream_ptr else None))

    def close(self):
        if self.h:
            self.lib.opf_ctx_destroy(self.h)
            self.h = None


# ---------------------------------------------------------------- mesh
def uniform_mesh_arrays(nnodes, lo, hi, pad=5):
    """Node coordinates/spacings of MeshBuilder::set1DMesh(min,max) + Symm ext."""
    lib = L.load()
    x = np.zeros(nnodes + 2 * pad)
    dx = np.zeros(nnodes + 2 * pad - 1)
    L.check(lib.opf_mesh_uniform_arrays(nnodes, lo, hi, pad, x.ctypes.data_as(L._DP), dx.ctypes.data_as(L._DP)))
    return x, dx


class CartesianMesh:
    def __init__(self, ctx, dims, ranges, pad=5):
        self.ctx = ctx
        self.dim = len(dims)
        self.dims = list(dims)
        self.pad = pad
        self.xs, self.dxs = [], []
        for d in range(self.dim):
            x, dx_ = uniform_mesh_arrays(dims[d], ranges[d][0], ranges[d][1], pad)
            self.xs.append(x)
            self.dxs.append(dx_)
        lib = ctx.lib
        xp = (L._DP * 3)()
        dp = (L._DP * 3)()
        for d in range(self.dim):
            xp[d] = self.xs[d].ctypes.data_as(L._DP)
            dp[d] = self.dxs[d].ctypes.data_as(L._DP)
        start = (C.c_int * 3)(0, 0, 0)
        nn = (C.c_int * 3)(*(self.dims + [1] * (3 - self.dim)))
        h = C.c_void_p()
        L.check(lib.opf_mesh_create(ctx.h, self.dim, start, nn, pad, xp, dp, C.byref(h)))
        self.h = h

    def x(self, d, i):
        return self.xs[d][i + self.pad]

    def dx(self, d, i):
        return self.dxs[d][i + self.pad]

    def getRange(self):
        return [(0, n) for n in self.dims]


class MeshBuilder:
    """MeshBuilder<CartesianMesh<Meta::int_<dim>>>().newMesh(...).setMeshOfDim(d, lo, hi).build()."""

    def __init__(self, ctx):
        self.ctx = ctx
        self._dims = None
        self._ranges = {}
        self._pad = 5

    def newMesh(self, *n):
        self._dims = list(n)
        return self

    def setMeshOfDim(self, d, lo, hi):
        self._ranges[d] = (lo, hi)
        return self

    def setPadWidth(self, p):
        self._pad = p
        return self

    def build(self):
        rng = [self._ranges.get(d, (0.0, 1.0)) for d in range(len(self._dims))]
        return CartesianMesh(self.ctx, self._dims, rng, self._pad)


# ---------------------------------------------------------------- expressions
_N_FIELD, _N_CONST, _N_ADD, _N_SUB, _N_MUL, _N_DIV, _N_NEG = 0, 1, 2, 3, 4, 5, 6
_N_D2C, _N_D1C, _N_D1UP, _N_D1DOWN, _N_WENO_UP, _N_WENO_DOWN = 7, 8, 9, 10, 11, 12
_N_INTP_C2N, _N_INTP_N2C, _N_UNKNOWN = 13, 14, 15
_N_FLUXLIM, _N_MIN, _N_MAX, _N_ABS, _N_GT, _N_LT, _N_GE, _N_LE, _N_COND = 16, 17, 18, 19, 20, 21, 22, 23, 24
_N_FUNC1, _N_FUNC2, _N_CONV = 25, 26, 27
# OPF_FN_* / OPF_FN2_* (the <cmath> names of Arithmetic/AMDS.hpp:39-90)
_FN1 = ["exp", "exp2", "expm1", "log", "log10", "log2", "log1p", "sqrt", "cbrt", "sin", "cos", "tan", "asin", "acos",
        "atan", "sinh", "cosh", "tanh", "asinh", "acosh", "atanh", "erf", "erfc", "tgamma", "lgamma", "ceil", "floor",
        "trunc", "round", "nearbyint", "rint", "logb"]
_FN2 = ["pow", "fmod", "remainder", "fdim", "hypot", "atan2", "copysign", "nextafter"]
_FN_SMOOTH_DELTA, _FN_SMOOTH_HEAVISIDE = 32, 33


def _wrap(v):
    if isinstance(v, Expr):
        return v
    if isinstance(v, (int, float, np.floating, np.integer)):
        return Expr(_N_CONST, value=float(v))
    raise TypeError(f"cannot use {type(v)} in an expression")


class Expr:
    """Lazy expression node (Expr/Expression.hpp); lowered by the library."""

    def __init__(self, kind, a=None, b=None, dim=0, value=0.0, field=None, c=None):
        self.kind, self.a, self.b, self.dim, self.value, self.field = kind, a, b, dim, value, field
        self.c = c  # conditional: the value selected when the condition fails

    def __add__(self, o): return Expr(_N_ADD, self, _wrap(o))
    def __radd__(self, o): return Expr(_N_ADD, _wrap(o), self)
    def __sub__(self, o): return Expr(_N_SUB, self, _wrap(o))
    def __rsub__(self, o): return Expr(_N_SUB, _wrap(o), self)
    def __mul__(self, o): return Expr(_N_MUL, self, _wrap(o))
    def __rmul__(self, o): return Expr(_N_MUL, _wrap(o), self)
    def __truediv__(self, o): return Expr(_N_DIV, self, _wrap(o))
    def __rtruediv__(self, o): return Expr(_N_DIV, _wrap(o), self)
    def __neg__(self): return Expr(_N_NEG, self)
    # comparisons (Operator/Logical/Compare.hpp) for conditional()
    def __gt__(self, o): return Expr(_N_GT, self, _wrap(o))
    def __lt__(self, o): return Expr(_N_LT, self, _wrap(o))
    def __ge__(self, o): return Expr(_N_GE, self, _wrap(o))
    def __le__(self, o): return Expr(_N_LE, self, _wrap(o))

    def __eq__(self, o):  # Equation: lhs == rhs (Equation/Equation.hpp:27-55)
        return Equation(self, _wrap(o))

    __hash__ = object.__hash__

    def flatten(self, unknown=None):
        """Post-order node list + field slot list for the C ABI."""
        nodes, fields, memo = [], [], {}
        slot_of = {}

        def visit(e):
            if id(e) in memo:
                return memo[id(e)]
            if e.kind == _N_UNKNOWN:  # a set's StencilField placeholder: a = the target's index
                node = (_N_UNKNOWN, int(e.a), 0, 0, 0.0)
            elif e.kind == _N_FIELD and unknown is not None and e.field is unknown:
                node = (_N_UNKNOWN, 0, 0, 0, 0.0)
            elif e.kind == _N_FIELD:
                key = id(e.field)
                if key not in slot_of:
                    slot_of[key] = len(fields)
                    fields.append(e.field)
                node = (_N_FIELD, slot_of[key], 0, 0, 0.0)
            elif e.kind == _N_CONST:
                node = (_N_CONST, 0, 0, 0, e.value)
            elif e.kind == _N_CONV:  # the weights follow the argument as consecutive CONST nodes
                ia = visit(e.a)
                ib = len(nodes)
                for w in e.value:
                    nodes.append((_N_CONST, 0, 0, 0, float(w)))
                node = (_N_CONV, ia, ib, e.dim, 0.0)
            else:
                ia = visit(e.a)
                ib = visit(e.b) if e.b is not None else 0
                dim = visit(e.c) if e.kind == _N_COND else e.dim
                node = (e.kind, ia, ib, dim, float(e.value) if e.kind in (_N_FLUXLIM, _N_FUNC1) else 0.0)
            nodes.append(node)
            memo[id(e)] = len(nodes) - 1
            return memo[id(e)]

        root = visit(self)
        return nodes, fields, root


class StencilUnknown(Expr):
    """The StencilField of target `index` of an equation set, as EqnHolder hands it to
    every getter (Equation/EquationHolder.hpp:31-99)."""

    def __init__(self, index, target):
        super().__init__(_N_UNKNOWN, a=int(index), field=target)


class Equation:
    """lhs == rhs (Equation/Equation.hpp:27-55). Python evaluates `1.0 == expr`
    as `expr == 1.0`, which swaps the sides; solves are unaffected (the operator's
    orientation is normalised) but the CSR form of a system is lhs - rhs, so
    build `Equation(1.0, expr)` to keep a constant on the left."""

    def __init__(self, lhs, rhs):
        self.lhs, self.rhs = _wrap(lhs), _wrap(rhs)


def _node_array(nodes):
    arr = (L.ExprNode * len(nodes))()
    for i, (k, a, b, d, v) in enumerate(nodes):
        arr[i].kind, arr[i].a, arr[i].b, arr[i].dim, arr[i].value = k, a, b, d, v
    return arr


# operator kernels (FDMOperators/*.hpp, Interpolator/D1Linear.hpp)
class _Kernel:
    kind = None


class D2SecondOrderCentered(_Kernel):
    kind = _N_D2C


class D1FirstOrderCentered(_Kernel):
    kind = _N_D1C


class D1FirstOrderBiasedUpwind(_Kernel):
    kind = _N_D1UP


class D1FirstOrderBiasedDownwind(_Kernel):
    kind = _N_D1DOWN


class D1WENO53Upwind(_Kernel):
    kind = _N_WENO_UP


class D1WENO53Downwind(_Kernel):
    kind = _N_WENO_DOWN


def d1(kernel, d, e):
    return Expr(kernel.kind, _wrap(e), dim=d)


def d2(kernel, d, e):
    return Expr(kernel.kind, _wrap(e), dim=d)


def d2x(kernel, e): return d2(kernel, 0, e)
def d2y(kernel, e): return d2(kernel, 1, e)
def d2z(kernel, e): return d2(kernel, 2, e)
def dx(kernel, e): return d1(kernel, 0, e)
def dy(kernel, e): return d1(kernel, 1, e)
def dz(kernel, e): return d1(kernel, 2, e)


class IntpDirection:  # Interpolator/IntpInterface.hpp:20
    Cor2Cen, Cen2Cor = 0, 1


class D1FluxLimiterGen:
    """D1FluxLimiterGen<Kernel>::Op<d, dir> (Interpolator/D1FluxLimiter.hpp:205-209): the kernel
    id of FluxLimiterKernels.hpp; the interpolation functions below take it as `kernel`."""

    def __init__(self, kernel_id):
        self.id = kernel_id


(D1Central, D1QUICK, D1CUI, D1Fromm, D1LinearUpwind, D1Minmod, D1Superbee, D1MUSCL, D1Harmonic, D1Albada,
 D1UMIST, D1SPL13, D1SPLMax12, D1SPLMax13, D1KorenLimitedCUI, D1SMART, D1BSOU) = [D1FluxLimiterGen(k) for k in range(17)]


def _fluxlim(d, u, e, kernel, direction):
    # value: kernel id + 32 x the named direction (1 Cen2Cor, 2 Cor2Cen), checked by the library
    return Expr(_N_FLUXLIM, _wrap(u), _wrap(e), dim=d,
                value=float(kernel.id + 32 * (1 if direction == IntpDirection.Cen2Cor else 2)))


def d1IntpCenterToCorner(d, *args, kernel=None):
    """IntpInterface.hpp:26-34: d1IntpCenterToCorner(d, e) is D1Linear<d, Cen2Cor>(e);
    d1IntpCenterToCorner(d, u, e, kernel=D1QUICK) the flux-limiter interpolation of e
    upwinded by u (D1FluxLimiter.hpp)."""
    if kernel is None:
        (e,) = args
        return Expr(_N_INTP_C2N, _wrap(e), dim=d)
    u, e = args
    return _fluxlim(d, u, e, kernel, IntpDirection.Cen2Cor)


def d1IntpCornerToCenter(d, *args, kernel=None):  # D1Linear<d, Cor2Cen>, or the limiter form
    if kernel is None:
        (e,) = args
        return Expr(_N_INTP_N2C, _wrap(e), dim=d)
    u, e = args
    return _fluxlim(d, u, e, kernel, IntpDirection.Cor2Cen)


def conditional(c, t, f):
    """conditional(c, t, f) (Operator/Conditional.hpp:181): t where c holds, else f."""
    return Expr(_N_COND, _wrap(c), _wrap(t), c=_wrap(f))


def max(a, b):  # noqa: A001  OpFlow::max (Operator/Arithmetic/MinMax.hpp): std::max elementwise
    return Expr(_N_MAX, _wrap(a), _wrap(b))


def min(a, b):  # noqa: A001  OpFlow::min: std::min elementwise
    return Expr(_N_MIN, _wrap(a), _wrap(b))


def _fn1_maker(fid, name):
    def f(e):
        return Expr(_N_FUNC1, _wrap(e), dim=fid)
    f.__name__ = name
    f.__doc__ = f"OpFlow::{name}(e): std::{name} of each value (Arithmetic/AMDS.hpp, DEFINE_UNIFUNC)."
    return f


def _fn2_maker(fid, name):
    def f(a, b):
        return Expr(_N_FUNC2, _wrap(a), _wrap(b), dim=fid)
    f.__name__ = name
    f.__doc__ = f"OpFlow::{name}(a, b): std::{name} of each pair of values (Arithmetic/AMDS.hpp, DEFINE_BINFUNC)."
    return f


for _i, _n in enumerate(_FN1):
    globals()[_n] = _fn1_maker(_i, _n)
for _i, _n in enumerate(_FN2):
    globals()[_n] = _fn2_maker(_i, _n)


def abs(e):  # noqa: A001  OpFlow::abs (Arithmetic/AMDS.hpp): std::abs of each value
    return Expr(_N_ABS, _wrap(e))


def smoothDelta(eps, e):
    """Math::smoothDelta(eps, d) of each value (Math/Function/Numeric.hpp:94-96), the functor the
    level-set example applies with UniOpAdaptor."""
    return Expr(_N_FUNC1, _wrap(e), dim=_FN_SMOOTH_DELTA, value=float(eps))


def smoothHeviside(eps, e):
    """Math::smoothHeviside(eps, d) of each value (Math/Function/Numeric.hpp:86-92)."""
    return Expr(_N_FUNC1, _wrap(e), dim=_FN_SMOOTH_HEAVISIDE, value=float(eps))


def conv(e, kernel):
    """conv(e, kernel) (Operator/Convolution/Convolution.hpp): sum over the window centred at
    each cell of kernel[offset] * e(cell + offset); kernel: array of odd extents, indexed
    [x, y(, z)] like DS::FixedSizeTensor<T, n0, n1(, n2)> (x-fastest values)."""
    k = np.asarray(kernel, dtype=np.float64)
    ext = list(k.shape) + [1] * (3 - k.ndim)
    if k.ndim < 1 or k.ndim > 3 or any(n % 2 == 0 for n in ext):
        raise ValueError("conv kernel: 1 to 3 dims of odd extents")
    weights = tuple(float(w) for w in k.reshape(-1, order="F"))  # x-fastest
    return Expr(_N_CONV, _wrap(e), dim=ext[0] | ext[1] << 8 | ext[2] << 16, value=weights)


# ---------------------------------------------------------------- fields
class CartesianField(Expr):
    """Device-resident CartesianField (CartesianField.hpp:37-794)."""

    def __init__(self, mesh, desc, name=""):
        super().__init__(_N_FIELD, field=self)
        self.mesh, self.ctx, self.name = mesh, mesh.ctx, name
        self.lib = mesh.ctx.lib
        h = C.c_void_p()
        L.check(self.lib.opf_field_create(mesh.h, C.byref(desc), C.byref(h)))
        self.h = h
        self.desc = L.FieldDesc.from_buffer_copy(self.lib.opf_field_get_desc(h).contents)
        self.dtype = np.float64 if desc.dtype == L.OPF_F64 else np.float32

    # --- ranges (DS::Range as (start, end) tuples per dim)
    def _rng(self, s, e):
        return [(getattr(self.desc, s)[d], getattr(self.desc, e)[d]) for d in range(self.desc.dim)]

    @property
    def localRange(self): return self._rng("local_start", "local_end")
    @property
    def assignableRange(self): return self._rng("assign_start", "assign_end")
    @property
    def accessibleRange(self): return self._rng("access_start", "access_end")
    @property
    def logicalRange(self): return self._rng("logical_start", "logical_end")

    @property
    def shape(self):
        """numpy shape of the storage block (z, y, x), x fastest."""
        e = self.desc.extent
        return tuple(e[d] for d in reversed(range(self.desc.dim)))

    def offset(self):
        return [self.desc.offset[d] for d in range(self.desc.dim)]

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        L.check(self.lib.opf_field_download(self.h, out.ctypes.data, out.nbytes))
        return out

    def upload(self, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype).reshape(self.shape)
        L.check(self.lib.opf_field_upload(self.h, a.ctypes.data, a.nbytes))

    def __getitem__(self, idx):
        """field[i] inside a rangeReduce functor: the field's value expression."""
        if isinstance(idx, SymIndex):
            return self
        raise TypeError("field[i] takes the symbolic index of rangeReduce; read values with numpy()")

    def index(self, *idx):
        """storage tuple (numpy index) of a field index (i0, i1, ...)."""
        return tuple(idx[d] - self.desc.offset[d] for d in reversed(range(self.desc.dim)))

    def updatePadding(self):
        L.check(self.lib.opf_field_update_padding(self.h))

    def assign(self, rhs, op=BasicArithOp.Eq):
        """OpFlow's `field (op)= rhs` (Expr::operator=, FieldAssigner::assign)."""
        if isinstance(rhs, (int, float, np.floating, np.integer)):
            L.check(self.lib.opf_field_assign_const(self.h, int(op), float(rhs)))
            return self
        nodes, fields, root = _wrap(rhs).flatten()
        arr = _node_array(nodes)
        fh = (C.c_void_p * builtins.max(1, len(fields)))(*[f.h.value for f in fields])
        L.check(self.lib.opf_sweep(arr, len(nodes), root, fh, len(fields), self.h, int(op)))
        return self

    def copy_from(self, other):
        L.check(self.lib.opf_field_copy(self.h, other.h))

    def clone(self, name=""):
        """Copy construction `auto q = p;` (CartesianField.hpp:57-68): same descriptor, copied data."""
        f = CartesianField(self.mesh, self.desc, name or self.name)
        f.copy_from(self)
        return f

    def cell_coords(self):
        """Coordinates initBy passes per dim over assignable∩local: x (corner) or x + .5 dx (centre)."""
        d = self.desc
        out = []
        for q in range(d.dim):
            s_, e_ = builtins.max(d.assign_start[q], d.local_start[q]), builtins.min(d.assign_end[q], d.local_end[q])
            idx = np.arange(s_, e_)
            x = self.mesh.xs[q][idx + self.mesh.pad]
            if d.loc[q] == LocOnMesh.Center:
                x = x + .5 * self.mesh.dxs[q][idx + self.mesh.pad]
            out.append((s_, e_, x))
        return out

    def initBy(self, fn):
        """initBy(f) (CartesianField.hpp:283-296): f(coords) over assignable∩local, then updatePadding.
        fn receives a list of broadcastable coordinate arrays (x0, x1[, x2]) and returns values."""
        d = self.desc
        cc = self.cell_coords()
        grids = np.meshgrid(*[c[2] for c in reversed(cc)], indexing="ij")[::-1]
        vals = np.asarray(fn(grids), dtype=np.float64)
        arr = self.numpy().astype(np.float64)
        sl = tuple(slice(cc[q][0] - d.offset[q], cc[q][1] - d.offset[q]) for q in reversed(range(d.dim)))
        arr[sl] = vals
        self.upload(arr)
        self.updatePadding()
        return self

    def ftcs(self, coef, nsteps):
        """Fused fast path of u = u + coef*(d2x(u)+d2y(u)[+d2z(u)]) repeated nsteps times."""
        L.check(self.lib.opf_ftcs(self.h, float(coef), int(nsteps)))

    def ftcs_blocked(self, coef, nsteps, tblock):
        L.check(self.lib.opf_ftcs_blocked(self.h, float(coef), int(nsteps), int(tblock)))

    def ftcs_plan(self, nsteps, tblock=0):
        """(kernel launches, most steps per launch) that ftcs/ftcs_blocked use for nsteps."""
        import ctypes as C
        nl, mx = C.c_int(), C.c_int()
        L.check(self.lib.opf_ftcs_plan(self.h, int(nsteps), int(tblock), C.byref(nl), C.byref(mx)))
        return nl.value, mx.value

    def device_ptr(self):
        return self.lib.opf_field_device_ptr(self.h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.opf_field_destroy(self.h)
            self.h = None


class ExprBuilder:
    """ExprBuilder<CartesianField>() (CartesianField.hpp:797-1033)."""

    def __init__(self):
        self.d = L.FieldDesc()
        self.mesh = None
        self.name = ""
        self.dtype = L.OPF_F64
        for i in range(3):
            self.d.loc[i] = LocOnMesh.Corner  # OpFlow's default location

    def setName(self, n):
        self.name = n
        return self

    def setMesh(self, m):
        self.mesh = m
        self.d.dim = m.dim
        return self

    def setLoc(self, loc):
        if isinstance(loc, (list, tuple)):
            for i, l_ in enumerate(loc):
                self.d.loc[i] = int(l_)
        else:
            for i in range(3):
                self.d.loc[i] = int(loc)
        return self

    def setLocOfDim(self, i, loc):
        self.d.loc[i] = int(loc)
        return self

    def setBC(self, d, pos, type_, val=0.0):
        self.d.bc_type[d][int(pos)] = int(type_)
        self.d.bc_value[d][int(pos)] = float(val)
        return self

    def setExt(self, *args):
        if len(args) == 1:
            for d in range(3):
                self.d.ext[d][0] = self.d.ext[d][1] = args[0]
        else:
            d, pos, w = args
            self.d.ext[d][int(pos)] = w
        return self

    def setPadding(self, p):
        self.d.padding = p
        return self

    def setDType(self, dt):
        self.dtype = L.OPF_F32 if dt in ("f32", np.float32, "float32") else L.OPF_F64
        return self

    def descriptor(self):
        d = L.FieldDesc.from_buffer_copy(self.d)
        d.dtype = self.dtype
        start = (C.c_int * 3)(0, 0, 0)
        nn = (C.c_int * 3)(*(self.mesh.dims + [1] * (3 - self.mesh.dim)))
        L.check(self.mesh.ctx.lib.opf_field_desc_finalize(C.byref(d), start, nn))
        return d

    def build(self):
        return CartesianField(self.mesh, self.descriptor(), self.name)


_RED = {"sum": 0, "max": 1, "min": 2, "sumsq": 3, "maxabs": 4}


class SymIndex:
    """The index a rangeReduce functor receives (RangeFor.hpp:86-121 passes each MDIndex of the
    range): a symbolic one, so `f[idx]` is the field's value expression and the functor's
    result is reduced on the device."""


def _reduction_kind(op):
    if isinstance(op, str):
        return op
    f = lambda a, b: float(op(a, b))  # noqa: E731
    if f(1.0, 2.0) == 3.0 and f(-4.0, 7.5) == 3.5 and f(0.25, 0.5) == 0.75:
        return "sum"
    if f(1.0, 2.0) == 2.0 and f(7.5, -4.0) == 7.5 and f(-3.0, -2.0) == -2.0:
        return "max"
    if f(1.0, 2.0) == 1.0 and f(7.5, -4.0) == -4.0 and f(-3.0, -2.0) == -3.0:
        return "min"
    raise L.OpfError(L.OPF_ERR_UNSUPPORTED, "rangeReduce: the reduction operator is not a sum, max or min")


def rangeReduce(a, b, c=None):
    """rangeReduce(range, op, func) (RangeFor.hpp:86-121), e.g.
        rangeReduce(p.assignableRange, lambda x, y: x + y, lambda i: p[i])
    op: a binary operator (sum, max or min, told from its values) or one of
    "sum", "max", "min", "sumsq", "maxabs"; func receives a symbolic index and returns a
    field or an expression of fields (evaluated into a temporary with the layout of the
    first field it reads). Deterministic on the device.
    Also rangeReduce(field, op[, range]) (range defaults to the assignable range)."""
    if isinstance(a, CartesianField):
        return _reduce_field(a, _reduction_kind(b), c or a.assignableRange)
    rng, red = a, _reduction_kind(b)
    e = _wrap(c(SymIndex()))
    if isinstance(e, CartesianField):
        return _reduce_field(e, red, rng)
    nodes, fields, _ = e.flatten()
    if not fields:
        raise L.OpfError(L.OPF_ERR_ARG, "rangeReduce: the functor reads no field")
    tmp = CartesianField(fields[0].mesh, fields[0].desc, "reduce_tmp")
    try:
        tmp.assign(e)
        return _reduce_field(tmp, red, rng)
    finally:
        tmp.close()


def _reduce_field(field, red, rng):
    dim = field.desc.dim
    rs = (C.c_int * 3)(*([r[0] for r in rng] + [0] * (3 - dim)))
    re_ = (C.c_int * 3)(*([r[1] for r in rng] + [1] * (3 - dim)))
    out = C.c_double()
    L.check(field.lib.opf_reduce(field.h, _RED[red], rs, re_, C.byref(out)))
    return out.value


def gen_bumps(field, seed, nbumps, noise=0.0):
    """Host array (storage layout) of seeded Gaussian bumps (+ noise) at cell centres."""
    out = np.zeros(field.shape, dtype=np.float64)
    L.check(field.lib.opf_gen_bumps(field.mesh.h, C.byref(field.desc), C.c_uint64(seed), nbumps, noise,
                                    out.ctypes.data_as(L._DP)))
    return out


def gen_uniform(seed, n, lo=0.0, hi=1.0):
    lib = L.load()
    out = np.zeros(n, dtype=np.float64)
    L.check(lib.opf_gen_uniform(C.c_uint64(seed), n, lo, hi, out.ctypes.data_as(L._DP)))
    return out


# ---------------------------------------------------------------- field streams
class TimeStamp(float):
    """Utils::TimeStamp: the time tag written with the next field."""


class RawBinaryOStream:
    """Utils::RawBinaryOStream (Utils/Writers/RawBinaryStream.hpp:42-60, 98-155):
    `os << TimeStamp(t) << u` writes <dir>/<name>_<count>.cart, dir being the given
    path's parent (std::filesystem::path::parent_path), values streamed from the GPU."""

    def __init__(self, path="./"):
        import os
        self.dir = os.path.dirname(str(path)) or "."
        self.time = 0.0
        self.count = 0

    def setCounterTo(self, c):
        self.count = int(c)

    def __lshift__(self, x):
        if isinstance(x, TimeStamp):
            self.time = float(x)
            return self
        L.check(x.lib.opf_field_write_cart(x.h, self.dir.encode(), x.name.encode(), self.count, self.time))
        self.count += 1
        return self


class RawBinaryIStream:
    """Utils::RawBinaryIStream (RawBinaryStream.hpp:62-77, 157-224): `is >> u` reads
    <dir>/<name>_<count>.cart into u (layout checked), then updatePadding."""

    def __init__(self, path="./"):
        import os
        self.dir = os.path.dirname(str(path)) or "."
        self.count = 0
        self.time = 0.0

    def setCounterTo(self, c):
        self.count = int(c)

    def __rshift__(self, f):
        t = C.c_double()
        L.check(f.lib.opf_field_read_cart(f.h, self.dir.encode(), f.name.encode(), self.count, C.byref(t)))
        self.time = t.value
        self.count += 1
        return self


StreamIn, StreamOut, StreamASCII, StreamBinary = 0x1, 0x2, 0x4, 0x8  # Core/Constants.hpp


class IOGroup:
    """Utils::IOGroup (Utils/Writers/IOGroup.hpp:23-167) over the raw binary streams: one
    stream per field at root + the field's name (files <dir of root>/<name>_<count>.cart);
    dump(t) writes every field stamped t (RawBinaryOStream), read(t) reads each field's
    next record (RawBinaryIStream: records are numbered, not indexed by time). One file
    per field: setAllInOne(True) with several fields is refused."""

    def __init__(self, stream, root, mode, *fields):
        self.stream_cls, self.root, self.mode, self.fields = stream, str(root), mode, list(fields)
        self.streams = None

    def _init(self):
        if self.streams is None:
            self.streams = [self.stream_cls(self.root + f.name) for f in self.fields]

    def dump(self, t):
        if self.stream_cls is not RawBinaryOStream:
            raise TypeError("IOGroup.dump needs an output stream")
        self._init()
        for s, f in zip(self.streams, self.fields):
            s << TimeStamp(float(t)) << f

    def read(self, t=None):
        if self.stream_cls is not RawBinaryIStream:
            raise TypeError("IOGroup.read needs an input stream")
        self._init()
        for s, f in zip(self.streams, self.fields):
            s >> f

    def setAllInOne(self, o):
        if o and len(self.fields) > 1:
            raise ValueError("IOGroup: the raw binary streams write one file per field")

    def fixedMesh(self):
        pass

    def dumpToSeparateFile(self):
        pass


def makeIOGroup(stream, root, *args):
    """Utils::makeIOGroup<Stream>(root, [mode,] fields...)"""
    if args and isinstance(args[0], int):
        return IOGroup(stream, root, args[0], *args[1:])
    return IOGroup(stream, root, StreamOut, *args)


class Utils:  # the reference's OpFlow::Utils namespace
    TimeStamp = TimeStamp
    RawBinaryOStream = RawBinaryOStream
    RawBinaryIStream = RawBinaryIStream
    IOGroup = IOGroup
    makeIOGroup = staticmethod(makeIOGroup)


from .solve import (StructSolverParams, EqnSolveState, Solve, makeEqnSolveHandler,  # noqa: E402
                    PrecondStructSolver, PoissonGMG, Server, DeviceBuffer, poisson_desc, gen_bumps_grid,
                    gen_bumps_grid_dev, SolveEqns, makeEqnHolder, makeStencilHolder, CSRMatrixGenerator,
                    ColoredMDRangeMapper, IJSolverParams)
