"""numpy restatement of OpFlow's lazy expression evaluation -- TEST ONLY.

evaluate() computes an expression tree (the post-order node list of the C
ABI) over an index box the way Expression::evalAtImpl_final recurses
(Expr/Expression.hpp:51-65): each operator evaluates its argument at shifted
indices and combines them with the mesh spacings in the reference's operation
order. numpy float64 elementwise ops round exactly like the reference's
scalar C++ (no FMA contraction), so results are comparable bit for bit.
"""
import math

import numpy as np

FIELD, CONST, ADD, SUB, MUL, DIV, NEG = 0, 1, 2, 3, 4, 5, 6
D2C, D1C, D1UP, D1DOWN, WENO_UP, WENO_DOWN, INTP_C2N, INTP_N2C, UNKNOWN = 7, 8, 9, 10, 11, 12, 13, 14, 15
FLUXLIM, MIN, MAX, ABS, GT, LT, GE, LE, COND = 16, 17, 18, 19, 20, 21, 22, 23, 24
FUNC1, FUNC2, CONV = 25, 26, 27
CORNER, CENTER = 0, 1
PI = 3.141592653589793  # Core/Constants.hpp:37
# functions of FUNC1 / FUNC2 (the <cmath> names of Arithmetic/AMDS.hpp:39-90)
FN1_NAMES = ["exp", "exp2", "expm1", "log", "log10", "log2", "log1p", "sqrt", "cbrt", "sin", "cos", "tan", "asin",
             "acos", "atan", "sinh", "cosh", "tanh", "asinh", "acosh", "atanh", "erf", "erfc", "tgamma", "lgamma",
             "ceil", "floor", "trunc", "round", "nearbyint", "rint", "logb", "smoothDelta", "smoothHeviside"]
FN2_NAMES = ["pow", "fmod", "remainder", "fdim", "hypot", "atan2", "copysign", "nextafter"]
FN_SMOOTH_DELTA, FN_SMOOTH_HEAVISIDE = 32, 33
# results the C library specifies exactly (the others are accurate to a few ulp, differently
# rounded by each library)
FN1_EXACT = {"sqrt", "ceil", "floor", "trunc", "round", "nearbyint", "rint", "logb"}
FN2_EXACT = {"fmod", "remainder", "fdim", "copysign", "nextafter"}


def _cround(x):  # std::round: halfway cases away from zero
    f = math.floor(x)
    if x >= 0:
        return f + 1.0 if x - f >= 0.5 else f
    c = math.ceil(x)
    return c - 1.0 if c - x >= 0.5 else c


def _logb(x):
    if x == 0:
        return -math.inf
    if math.isinf(x):
        return math.inf
    if math.isnan(x):
        return x
    return float(math.frexp(x)[1] - 1)


def _fdim(x, y):
    if math.isnan(x) or math.isnan(y):
        return math.nan
    return x - y if x > y else 0.0


_FN1 = {"exp": math.exp, "exp2": lambda x: 2.0 ** x, "expm1": math.expm1, "log": math.log, "log10": math.log10,
        "log2": math.log2, "log1p": math.log1p, "sqrt": math.sqrt, "cbrt": lambda x: float(np.cbrt(x)),
        "sin": math.sin, "cos": math.cos, "tan": math.tan, "asin": math.asin, "acos": math.acos, "atan": math.atan,
        "sinh": math.sinh, "cosh": math.cosh, "tanh": math.tanh, "asinh": math.asinh, "acosh": math.acosh,
        "atanh": math.atanh, "erf": math.erf, "erfc": math.erfc, "tgamma": math.gamma, "lgamma": math.lgamma,
        "ceil": lambda x: float(math.ceil(x)), "floor": lambda x: float(math.floor(x)),
        "trunc": lambda x: float(math.trunc(x)), "round": _cround, "nearbyint": lambda x: float(np.rint(x)),
        "rint": lambda x: float(np.rint(x)), "logb": _logb}
_FN2 = {"pow": math.pow, "fmod": math.fmod, "remainder": math.remainder, "fdim": _fdim, "hypot": math.hypot,
        "atan2": math.atan2, "copysign": math.copysign, "nextafter": math.nextafter}


def fn1(fid, x, eps=0.0):
    """The <cmath> function fid of FUNC1 elementwise through Python's math module (the C
    library's functions); smoothDelta / smoothHeviside restate Math/Function/Numeric.hpp:86-96."""
    x = np.asarray(x, dtype=np.float64)
    if fid == FN_SMOOTH_DELTA:  # std::max(d, -d) > eps ? 0. : 1. / 2. / eps * (1 + std::cos(d * PI / eps))
        c = np.vectorize(math.cos, otypes=[np.float64])(x * PI / eps) if x.size else x
        return np.where(_max(x, -x) > eps, 0.0, 1. / 2. / eps * (1 + c))
    if fid == FN_SMOOTH_HEAVISIDE:  # d < -eps ? 0 : d > eps ? 1 : 0.5 * (1 + d / eps + 1 / PI * sin(d * PI / eps))
        sn = np.vectorize(math.sin, otypes=[np.float64])(x * PI / eps) if x.size else x
        return np.where(x < -eps, 0.0, np.where(x > eps, 1.0, 0.5 * (1 + x / eps + 1 / PI * sn)))
    f = _FN1[FN1_NAMES[fid]]
    return np.vectorize(f, otypes=[np.float64])(x)


def fn2(fid, x, y):
    x, y = np.broadcast_arrays(np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64))
    return np.vectorize(_FN2[FN2_NAMES[fid]], otypes=[np.float64])(x, y)
# limiter kernels of FLUXLIM (Interpolator/FluxLimiterKernels.hpp)
(LIM_CDS, LIM_QUICK, LIM_CUI, LIM_FROMM, LIM_LUI, LIM_MINMOD, LIM_SUPERBEE, LIM_MUSCL, LIM_HARMONIC, LIM_VANALBADA,
 LIM_UMIST, LIM_SPL13, LIM_SPLMAX12, LIM_SPLMAX13, LIM_KOREN, LIM_SMART, LIM_BSOU) = range(17)


def _min(a, b):  # std::min(a, b): (b < a) ? b : a
    return np.where(b < a, b, a)


def _max(a, b):  # std::max(a, b): (a < b) ? b : a
    return np.where(a < b, b, a)


def _ratio(num, den):
    """std::ratio<num, den> read as `num / den` with integer operands (FluxLimiterKernels.hpp:76-77)."""
    return float(int(num / den) if num * den >= 0 else -(abs(num) // abs(den)))


# (M, Alpha, Kappa) of the SPL / GPL kernels (FluxLimiterKernels.hpp:100-143)
_SPL = {LIM_UMIST: ((2, 1), (1, 2)), LIM_SPL13: ((2, 1), (1, 3)),
        LIM_SPLMAX12: ((2, 1), (1, 2)), LIM_SPLMAX13: ((2, 1), (1, 3))}
_GPL = {LIM_KOREN: ((2, 1), (0, 1), (1, 3)), LIM_SMART: ((4, 1), (0, 1), (1, 2)), LIM_BSOU: ((1, 1), (0, 1), (-1, 1))}
_KAPPA = {LIM_CDS: 1.0, LIM_QUICK: 0.5, LIM_CUI: 1. / 3., LIM_FROMM: 0.0, LIM_LUI: -1.0}


def limiter(kind, r):
    """psi(r) of the ratio-form kernels (FluxLimiterKernels.hpp:46-143), numpy, same operation order."""
    z = np.zeros_like(r)
    if kind == LIM_MINMOD:
        return _max(z, _min(r, 1.0 + z))
    if kind == LIM_SUPERBEE:
        return _max(_max(z, _min(2. * r, 1.0 + z)), _min(r, 2.0 + z))
    if kind == LIM_MUSCL:
        return _max(z, _min(_min(2 * r, (r + 1) / 2.), 2.0 + z))
    if kind == LIM_HARMONIC:
        return (r + np.fabs(r)) / (r + 1)
    if kind == LIM_VANALBADA:
        return r * (r + 1) / (r * r + 1)
    if kind in _SPL:
        (mn, md), (kn, kd) = _SPL[kind]
        m, kappa = _ratio(mn, md), _ratio(kn, kd)
        a = 0.5 * (1 + kappa) * r + 0.5 * (1 - kappa)
        b = 0.5 * (1 - kappa) * r + 0.5 * (1 + kappa)
        if kind in (LIM_UMIST, LIM_SPL13):
            inner = _min(_min(_min(m * r, a), b), m + z)
        else:
            inner = _min(_min(m * r, _max(a, b)), m + z)
        return _max(z, inner)
    (mn, md), (an, ad), (kn, kd) = _GPL[kind]
    m, alpha, kappa = _ratio(mn, md), _ratio(an, ad), _ratio(kn, kd)
    a = 0.5 * (1 + kappa) * r + 0.5 * (1 - kappa)
    return _max(z, _min(_min((2 + alpha) * r, a), m + z))


def fluxlim(kind, c2n, up, X, DXf, y, dxi):
    """One branch of D1FluxLimiterImpl::eval (D1FluxLimiter.hpp:43-129): X(k) the node coordinate at
    i + k, DXf(k) the spacing at i + k, y(k) the argument at i + k, dxi the spacing multiplying K."""
    if c2n:
        xc = lambda k: X(k) + 0.5 * DXf(k)  # noqa: E731  (cell centres)
        k1 = -2 if up else -1
    else:
        xc = X
        k1 = -1 if up else 0
    x1, x2, x3 = xc(k1), xc(k1 + 1), xc(k1 + 2)
    y1, y2, y3 = y(k1), y(k1 + 1), y(k1 + 2)
    if up:
        slop_u, slop_f = (y2 - y1) / (x2 - x1), (y3 - y2) / (x3 - x2)
    else:
        slop_u, slop_f = (y3 - y2) / (x3 - x2), (y2 - y1) / (x2 - x1)
    if kind in _KAPPA:
        kappa = _KAPPA[kind]
        k = (1 + kappa) / 2. * slop_f + (1 - kappa) / 2. * slop_u
        t = dxi * 0.5 * k
    else:
        r = slop_f / (slop_u + 1e-16)
        t = dxi * 0.5 * limiter(kind, r) * slop_u
    return y2 + t if up else y2 - t


def _locs(nodes, fields, dim):
    locs = []
    for (k, a, b, d, v) in nodes:
        if k == FIELD:
            desc = fields[a][1]
            locs.append([desc.loc[q] for q in range(dim)])
        elif k == CONST:
            locs.append(None)
        elif k in (ADD, SUB, MUL, DIV, MIN, MAX, GT, LT, GE, LE):
            locs.append(list(locs[a]) if locs[a] is not None else (list(locs[b]) if locs[b] is not None else None))
        elif k == COND:  # Conditional.hpp prepare: the properties of the selected value
            locs.append(list(locs[b]) if locs[b] is not None else (list(locs[d]) if locs[d] is not None else None))
        elif k in (ABS, FUNC1, CONV):
            locs.append(list(locs[a]) if locs[a] is not None else None)
        elif k == FUNC2:
            locs.append(list(locs[a]) if locs[a] is not None else (list(locs[b]) if locs[b] is not None else None))
        elif k == FLUXLIM:  # D1FluxLimiter.hpp prepare: the location of the interpolated argument, flipped
            l_ = list(locs[b]) if locs[b] is not None else None
            if l_ is not None:
                l_[d] = CORNER if l_[d] == CENTER else CENTER
            locs.append(l_)
        else:
            l_ = list(locs[a]) if locs[a] is not None else None
            if l_ is not None:
                if k == D1C:
                    l_[d] = CORNER if l_[d] == CENTER else CENTER
                if k == INTP_C2N:
                    l_[d] = CORNER
                if k == INTP_N2C:
                    l_[d] = CENTER
            locs.append(l_)
    return locs


def _weno_parts(v, h, up):
    """D1WENO53Upwind::kernel / D1WENO53Downwind::kernel (D1WENO53Upwind.hpp:133-151,
    D1WENO53Downwind.hpp:136-154); v = six values in index order. Returns the three
    candidate derivatives, the three smoothness indicators and the three weights."""
    if up:
        pm2, pm1, p0, pp1, pp2, pp3 = v
        d1, d2, d3, d4, d5 = (pp3 - pp2) / h, (pp2 - pp1) / h, (pp1 - p0) / h, (p0 - pm1) / h, (pm1 - pm2) / h
    else:
        pm3, pm2, pm1, p0, pp1, pp2 = v
        d1, d2, d3, d4, d5 = (pm2 - pm3) / h, (pm1 - pm2) / h, (p0 - pm1) / h, (pp1 - p0) / h, (pp2 - pp1) / h
    ddx1 = d1 / 3. - 7. * d2 / 6. + 11. * d3 / 6.
    ddx2 = -d2 / 6. + 5 * d3 / 6 + d4 / 3
    ddx3 = d3 / 3 + 5 * d4 / 6 - d5 / 6
    def pow2(t):  # Math::pow2 (Math/Function/Numeric.hpp:104)
        return t * t
    s1 = 13. / 12. * pow2(d1 - 2 * d2 + d3) + pow2(d1 - 4 * d2 + 3 * d3) / 4
    s2 = 13. / 12. * pow2(d2 - 2 * d3 + d4) + pow2(d2 - d4) / 4.
    s3 = 13. / 12. * pow2(d3 - 2 * d4 + d5) + pow2(3 * d3 - 4 * d4 + d5) / 4.
    m = np.maximum(np.maximum(np.maximum(np.maximum(d1 * d1, d2 * d2), d3 * d3), d4 * d4), d5 * d5)
    eps = 1e-6 * m + 1e-99
    a1 = .1 / pow2(s1 + eps)
    a2 = .6 / pow2(s2 + eps)
    a3 = .3 / pow2(s3 + eps)
    w1 = a1 / (a1 + a2 + a3)
    w2 = a2 / (a1 + a2 + a3)
    w3 = a3 / (a1 + a2 + a3)
    return (ddx1, ddx2, ddx3), (s1, s2, s3), (w1, w2, w3)


def _weno(v, h, up):
    """The WENO5 derivative: the weighted sum of the candidates (same operation order)."""
    (c1, c2, c3), _, (w1, w2, w3) = _weno_parts(v, h, up)
    return w1 * c1 + w2 * c2 + w3 * c3


def evaluate(nodes, root, fields, mesh, lo, hi):
    """Value of node `root` over the index box [lo, hi) (lists per dim).

    fields: list of (padded numpy array in (z,y,x) order, FieldDesc) per slot.
    Returns a float64 array of shape (hi-lo) reversed (z,y,x)."""
    dim = mesh.dim
    locs = _locs(nodes, fields, dim)

    def mesh_col(d, arr1d):
        shape = [1] * dim
        shape[dim - 1 - d] = arr1d.shape[0]
        return arr1d.reshape(shape)

    def DX(d, lo_, hi_, k):
        return mesh_col(d, mesh.dx(d, np.arange(lo_[d], hi_[d]) + k))

    def X(d, lo_, hi_, k):
        return mesh_col(d, mesh.x(d, np.arange(lo_[d], hi_[d]) + k))

    def sh(lo_, hi_, d, k):
        l2, h2 = list(lo_), list(hi_)
        l2[d] += k
        h2[d] += k
        return l2, h2

    def ev(n, lo_, hi_):
        k, a, b, d, v = nodes[n]
        if k == FIELD:
            arr, desc = fields[a]
            sl = tuple(slice(lo_[q] - desc.offset[q], hi_[q] - desc.offset[q]) for q in reversed(range(dim)))
            for q in range(dim):
                if lo_[q] < desc.offset[q] or hi_[q] > desc.offset[q] + desc.extent[q]:
                    raise IndexError(f"oracle read of slot {a} outside storage in dim {q}")
            return np.asarray(arr, dtype=np.float64)[sl]
        if k == CONST:
            return np.float64(v)
        if k == ADD:
            return ev(a, lo_, hi_) + ev(b, lo_, hi_)
        if k == SUB:
            return ev(a, lo_, hi_) - ev(b, lo_, hi_)
        if k == MUL:
            return ev(a, lo_, hi_) * ev(b, lo_, hi_)
        if k == DIV:
            return ev(a, lo_, hi_) / ev(b, lo_, hi_)
        if k == NEG:
            return -ev(a, lo_, hi_)
        if k in (MIN, MAX):
            x_, y_ = np.asarray(ev(a, lo_, hi_), dtype=np.float64), np.asarray(ev(b, lo_, hi_), dtype=np.float64)
            return _min(x_, y_) if k == MIN else _max(x_, y_)
        if k == ABS:
            return np.fabs(ev(a, lo_, hi_))
        if k == FUNC1:
            return fn1(d, ev(a, lo_, hi_), v)
        if k == FUNC2:
            return fn2(d, ev(a, lo_, hi_), ev(b, lo_, hi_))
        if k == CONV:  # Convolution::eval: rangeReduce_s x-fastest of kernel[idx - i + n/2] * e(idx)
            ks = [d & 255, (d >> 8) & 255, (d >> 16) & 255]
            res, q = None, 0
            for c2 in range(ks[2]):
                for c1 in range(ks[1]):
                    for c0 in range(ks[0]):
                        off = [c0 - ks[0] // 2, c1 - ks[1] // 2, c2 - ks[2] // 2]
                        l2 = [lo_[t] + off[t] for t in range(dim)]
                        h2 = [hi_[t] + off[t] for t in range(dim)]
                        term = np.float64(nodes[b + q][4]) * ev(a, l2, h2)
                        res = term if res is None else res + term
                        q += 1
            return res
        if k in (GT, LT, GE, LE):
            x_, y_ = ev(a, lo_, hi_), ev(b, lo_, hi_)
            c_ = {GT: np.greater, LT: np.less, GE: np.greater_equal, LE: np.less_equal}[k](x_, y_)
            return c_.astype(np.float64)
        if k == COND:  # CondOp::eval (Conditional.hpp:35-40): c ? t : f
            return np.where(np.asarray(ev(a, lo_, hi_)) != 0, ev(b, lo_, hi_), ev(d, lo_, hi_))
        if k == FLUXLIM:  # D1FluxLimiterImpl::eval: u[i] > 0. ? upwind : downwind
            c2n = locs[b] is not None and locs[b][d] == CENTER
            kind = int(v) % 32
            y = lambda kk: np.asarray(ev(b, *sh(lo_, hi_, d, kk)), dtype=np.float64)  # noqa: E731
            Xf = lambda kk: X(d, lo_, hi_, kk)  # noqa: E731
            Df = lambda kk: DX(d, lo_, hi_, kk)  # noqa: E731
            kd_up = -1 if c2n else 0
            up = fluxlim(kind, c2n, True, Xf, Df, y, DX(d, lo_, hi_, kd_up))
            dn = fluxlim(kind, c2n, False, Xf, Df, y, DX(d, lo_, hi_, 0))
            uu = np.asarray(ev(a, lo_, hi_), dtype=np.float64)
            return np.where(uu > 0., up, dn)
        corner = locs[a] is not None and locs[a][d] == CORNER
        if k == D2C:  # D2SecondOrderCentered::eval (D2SecondOrderCentered.hpp:160-171)
            l_ = ev(a, *sh(lo_, hi_, d, -1))
            c = ev(a, lo_, hi_)
            r = ev(a, *sh(lo_, hi_, d, +1))
            dxl = DX(d, lo_, hi_, -1) if corner else (DX(d, lo_, hi_, -1) + DX(d, lo_, hi_, 0)) * 0.5
            dxr = DX(d, lo_, hi_, 0) if corner else (DX(d, lo_, hi_, 0) + DX(d, lo_, hi_, 1)) * 0.5
            dxc = (dxl + dxr) * 0.5
            return ((r - c) / dxr - (c - l_) / dxl) / dxc
        if k == D1C:  # D1FirstOrderCentered::eval (D1FirstOrderCentered.hpp:30-37)
            if not corner:
                return (ev(a, lo_, hi_) - ev(a, *sh(lo_, hi_, d, -1))) / (DX(d, lo_, hi_, -1) + DX(d, lo_, hi_, 0)) * 2
            return (ev(a, *sh(lo_, hi_, d, 1)) - ev(a, lo_, hi_)) / DX(d, lo_, hi_, 0)
        if k == D1UP:  # D1FirstOrderBiasedUpwind::eval (D1FirstOrderBiasedUpwind.hpp:53-58)
            den = DX(d, lo_, hi_, 0) if corner else (DX(d, lo_, hi_, 0) + DX(d, lo_, hi_, 1)) * 0.5
            return (ev(a, *sh(lo_, hi_, d, 1)) - ev(a, lo_, hi_)) / den
        if k == D1DOWN:  # D1FirstOrderBiasedDownwind::eval (D1FirstOrderBiasedDownwind.hpp:52-57)
            den = DX(d, lo_, hi_, -1) if corner else (DX(d, lo_, hi_, -1) + DX(d, lo_, hi_, 0)) * 0.5
            return (ev(a, lo_, hi_) - ev(a, *sh(lo_, hi_, d, -1))) / den
        if k in (WENO_UP, WENO_DOWN):
            k0 = -2 if k == WENO_UP else -3
            vals = [ev(a, *sh(lo_, hi_, d, k0 + t)) for t in range(6)]
            return _weno(vals, DX(d, lo_, hi_, 0), k == WENO_UP)
        if k == INTP_C2N:  # D1Linear<d,Cen2Cor>::eval + Interpolator1D::intp (D1Linear.hpp:34-42)
            x1 = X(d, lo_, hi_, -1) + 0.5 * DX(d, lo_, hi_, -1)
            x2 = X(d, lo_, hi_, 0) + 0.5 * DX(d, lo_, hi_, 0)
            x = X(d, lo_, hi_, 0)
            y1 = ev(a, *sh(lo_, hi_, d, -1))
            y2 = ev(a, lo_, hi_)
            A, B = x1 - x, x2 - x
            return (A * y2 - B * y1) / (A - B)
        if k == INTP_N2C:  # D1Linear<d,Cor2Cen>: Math::mid (Interpolator.hpp:83)
            return (ev(a, lo_, hi_) + ev(a, *sh(lo_, hi_, d, 1))) * 0.5
        raise ValueError(f"node kind {k}")

    out = ev(root, list(lo), list(hi))
    shape = tuple(hi[q] - lo[q] for q in reversed(range(dim)))
    return np.broadcast_to(out, shape).astype(np.float64)


def assign(nodes, root, fields, mesh, out_arr, out_desc, op=0, update_padding=None):
    """OpFlow's `out (op)= expr`: evaluate over assignable∩local of out, write, then
    updatePadding via the given callable (FieldAssigner.hpp:26-86)."""
    dim = mesh.dim
    lo = [max(out_desc.assign_start[q], out_desc.local_start[q]) for q in range(dim)]
    hi = [min(out_desc.assign_end[q], out_desc.local_end[q]) for q in range(dim)]
    val = evaluate(nodes, root, fields, mesh, lo, hi)
    res = np.array(out_arr, dtype=np.float64, copy=True)
    sl = tuple(slice(lo[q] - out_desc.offset[q], hi[q] - out_desc.offset[q]) for q in reversed(range(dim)))
    if op == 0:
        res[sl] = val
    elif op == 1:
        res[sl] = res[sl] + val
    elif op == 2:
        res[sl] = res[sl] - val
    elif op == 3:
        res[sl] = res[sl] * val
    elif op == 4:
        res[sl] = res[sl] / val
    if update_padding is not None:
        res = update_padding(res)
    return res


def linear_system(eqs, unknowns, mesh, fill, pins=None):
    """Dense matrix and right-hand side of a set of equations that are linear in their
    unknowns, assembled the way CSRMatrixGenerator::generate does
    (Equation/CSRMatrixGenerator.hpp:26-41, 143-193) -- TEST ONLY, small grids.

    The reference builds row r of block i from the StencilPad of lhs_i - rhs_i at cell
    r (StencilField.hpp:123-350, StencilPad.hpp:136-331): an affine function of the
    unknowns whose ghost entries follow the boundary rule (ghost = alpha * mirror +
    beta). Here the same affine map is taken by linearity instead of pad algebra:
    with F(u) = evaluate(lhs - rhs) over fields whose interiors are u and whose
    ghosts are filled by updatePadding, bias = F(0) and column c = F(e_c) - F(0).
    Blocks follow ColoredMDRangeMapper (ColoredMDRangeMapper.hpp:20-41): block i is
    target i's cells x-fastest; pins[i] makes block i's last row x = 0 (rhs 0).

    eqs: list of (nodes, lhs_root, rhs_root, fields) with UNKNOWN nodes whose `a` is
    the target index; fields as for evaluate(). unknowns: list of (padded array,
    desc) per target (one layout). fill(i, arr) returns arr with target i's ghosts
    filled. Returns (A, rhs) with rhs = -bias."""
    neq = len(unknowns)
    arr0, desc = unknowns[0]
    dim = mesh.dim
    lo = [max(desc.assign_start[q], desc.local_start[q]) for q in range(dim)]
    hi = [min(desc.assign_end[q], desc.local_end[q]) for q in range(dim)]
    shape = tuple(hi[q] - lo[q] for q in reversed(range(dim)))
    N = int(np.prod(shape))
    sl = tuple(slice(lo[q] - desc.offset[q], hi[q] - desc.offset[q]) for q in reversed(range(dim)))

    def field_with(i, interior):
        a = np.zeros_like(np.asarray(arr0, dtype=np.float64))
        a[sl] = interior.reshape(shape)
        return fill(i, a)

    zero = [field_with(j, np.zeros(N)) for j in range(neq)]
    A = np.zeros((neq * N, neq * N))
    rhs = np.zeros(neq * N)
    for i, (nodes, lr, rr, fields) in enumerate(eqs):
        nf = len(fields)
        sub = [(FIELD, nf + a, 0, 0, 0.0) if k == UNKNOWN else (k, a, b, d, v) for (k, a, b, d, v) in nodes]
        sub.append((SUB, lr, rr, 0, 0.0))
        root = len(sub) - 1

        def F(us):
            return evaluate(sub, root, list(fields) + [(u, unknowns[j][1]) for j, u in enumerate(us)], mesh, lo,
                            hi).reshape(-1)

        f0 = F(zero)
        rhs[i * N:(i + 1) * N] = -f0
        reads = {a for (k, a, b, d, v) in nodes if k == UNKNOWN}
        for j in sorted(reads):
            for c in range(N):
                e = np.zeros(N)
                e[c] = 1.0
                us = list(zero)
                us[j] = field_with(j, e)
                A[i * N:(i + 1) * N, j * N + c] = F(us) - f0
        if pins and pins[i]:
            r = i * N + N - 1
            A[r, :] = 0.0
            A[r, r] = 1.0
            rhs[r] = 0.0
    return A, rhs
