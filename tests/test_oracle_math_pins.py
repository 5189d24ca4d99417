"""Pins of the oracle's math-function and convolution nodes (oracle/expr.py; CPU only).

  * conv: Convolution::eval (Convolution.hpp:52-68) sums kernel[idx - i + n/2] * e(idx) over
    the window with rangeReduce_s (RangeFor.hpp:50-60: the first term, then each next one
    added, the index advancing x-fastest as RangedIndex::operator++ does), and
    FixedSizeTensor stores its values x-fastest (FixedSizeTensor.hpp:63-66). Pinned by a
    direct double loop written from those two facts, on an asymmetric kernel (a transposed
    kernel or a reversed window fails it) and by the closed forms of a uniform field and of
    linear data (the level-set kernel of weights summing to one preserves both);
  * the unary / binary functions: the C library through Python's math module, checked
    against numpy's correctly rounded sqrt and against identities that fix the function
    id mapping (exp(log x) = x, sin^2 + cos^2 = 1, hypot(3, 4) = 5, ...);
  * smoothDelta / smoothHeviside: Math/Function/Numeric.hpp:86-96, pinned by their closed
    values (delta(0) = 1/eps, delta(+-eps) = 0 edge, heaviside(0) = 1/2, the integral of
    delta over [-eps, eps] = 1, heaviside' = delta).
"""
import math
from types import SimpleNamespace

import numpy as np
import pytest

import opflow_amd as of
from oracle import expr as E

PAD = 2


class Mesh2:
    dim = 2

    def x(self, d, i):
        return np.asarray(i, dtype=np.float64) * 0.1

    def dx(self, d, i):
        return np.full(np.shape(i), 0.1)


def _conv(vals, kernel2d):
    """evaluate() of CONV(FIELD) with the kernel given as a [x][y] array (FixedSizeTensor<3, 3>)."""
    ny, nx = vals.shape
    desc = SimpleNamespace(offset=[-PAD, -PAD], extent=[nx, ny], loc=[E.CENTER, E.CENTER])
    w = np.asarray(kernel2d, dtype=np.float64)
    flat = w.reshape(-1, order="F")  # x-fastest
    nodes = [(E.FIELD, 0, 0, 0, 0.0)] + [(E.CONST, 0, 0, 0, float(v)) for v in flat]
    nodes.append((E.CONV, 0, 1, w.shape[0] | w.shape[1] << 8 | 1 << 16, 0.0))
    return E.evaluate(nodes, len(nodes) - 1, [(vals, desc)], Mesh2(), [0, 0], [nx - 2 * PAD, ny - 2 * PAD])


def test_conv_direct_loop():
    rng = np.random.default_rng(4)
    vals = rng.normal(size=(12, 15))  # (y, x) storage from index -PAD
    ker = rng.normal(size=(3, 3))     # ker[x][y]
    got = _conv(vals, ker)
    ny, nx = got.shape
    want = np.zeros_like(got)
    for j in range(ny):
        for i in range(nx):
            s = None
            for b in range(3):          # y outer
                for a in range(3):      # x fastest
                    t = ker[a][b] * vals[j + b - 1 + PAD, i + a - 1 + PAD]
                    s = t if s is None else s + t
            want[j, i] = s
    np.testing.assert_array_equal(got, want)
    # a transposed kernel gives something else
    assert not np.array_equal(got, _conv(vals, ker.T))


def test_conv_level_set_kernel_closed_forms():
    c_, o_ = 16. / 24., 1. / 24.
    ker = [[o_, o_, o_], [o_, c_, o_], [o_, o_, o_]]
    flat = np.full((10, 11), 0.75)
    np.testing.assert_allclose(_conv(flat, ker), 0.75, rtol=1e-15)
    yy, xx = np.mgrid[0:10, 0:11].astype(np.float64)
    lin = 0.3 + 0.7 * xx - 0.2 * yy  # a symmetric kernel of unit sum reproduces linear data
    got = _conv(lin, ker)
    np.testing.assert_allclose(got, lin[PAD:-PAD, PAD:-PAD], rtol=0, atol=1e-14)


def test_python_conv_weights_x_fastest():
    k = np.arange(9.0).reshape(3, 3)  # k[x][y]
    e = of.conv(of.Expr(0), k)
    assert e.dim == 3 | 3 << 8 | 1 << 16
    assert e.value == tuple(k[x][y] for y in range(3) for x in range(3))


def test_fn1_exact_and_identities():
    x = np.linspace(0.1, 3.0, 97)
    ix = {n: i for i, n in enumerate(E.FN1_NAMES)}
    np.testing.assert_array_equal(E.fn1(ix["sqrt"], x), np.sqrt(x))
    np.testing.assert_allclose(E.fn1(ix["exp"], E.fn1(ix["log"], x)), x, rtol=4e-16)
    np.testing.assert_allclose(E.fn1(ix["sin"], x) ** 2 + E.fn1(ix["cos"], x) ** 2, 1.0, rtol=4e-16)
    np.testing.assert_allclose(E.fn1(ix["tanh"], x), np.tanh(x), rtol=4e-16)
    np.testing.assert_allclose(E.fn1(ix["log2"], 2.0 ** x), x, rtol=4e-16, atol=4e-16)
    np.testing.assert_allclose(E.fn1(ix["tgamma"], np.array([1.0, 2.0, 3.0, 4.0, 5.0])), [1, 1, 2, 6, 24])
    np.testing.assert_array_equal(E.fn1(ix["round"], np.array([-2.5, -1.5, -0.5, 0.5, 1.5, 2.5, 0.49999999999999994])),
                                  [-3.0, -2.0, -1.0, 1.0, 2.0, 3.0, 0.0])  # halfway away from zero
    np.testing.assert_array_equal(E.fn1(ix["rint"], np.array([-2.5, -1.5, 0.5, 1.5, 2.5])),
                                  [-2.0, -2.0, 0.0, 2.0, 2.0])  # halfway to even
    np.testing.assert_array_equal(E.fn1(ix["logb"], np.array([1.0, 3.0, 0.75, 1024.0])), [0.0, 1.0, -1.0, 10.0])
    np.testing.assert_array_equal(E.fn1(ix["trunc"], np.array([-1.7, 1.7])), [-1.0, 1.0])


def test_fn2_identities():
    ix = {n: i for i, n in enumerate(E.FN2_NAMES)}
    assert E.fn2(ix["hypot"], 3.0, 4.0) == 5.0
    assert E.fn2(ix["pow"], 2.0, 10.0) == 1024.0
    assert E.fn2(ix["fmod"], 7.5, 2.0) == 1.5
    assert E.fn2(ix["remainder"], 7.5, 2.0) == -0.5
    assert E.fn2(ix["fdim"], 3.0, 5.0) == 0.0 and E.fn2(ix["fdim"], 5.0, 3.0) == 2.0
    assert E.fn2(ix["atan2"], 1.0, 1.0) == math.pi / 4
    assert E.fn2(ix["copysign"], 2.0, -0.0) == -2.0
    assert E.fn2(ix["nextafter"], 1.0, 2.0) == 1.0 + 2.0 ** -52


def test_smooth_delta_heaviside_closed_forms():
    eps = 0.25
    D, H = E.FN_SMOOTH_DELTA, E.FN_SMOOTH_HEAVISIDE
    assert E.fn1(D, np.array([0.0]), eps)[0] == 1.0 / eps
    np.testing.assert_allclose(E.fn1(D, np.array([eps, -eps]), eps), 0.0, atol=1e-15)
    np.testing.assert_array_equal(E.fn1(D, np.array([0.3, -0.26]), eps), 0.0)
    assert E.fn1(H, np.array([0.0]), eps)[0] == 0.5
    np.testing.assert_array_equal(E.fn1(H, np.array([-0.3, 0.3]), eps), [0.0, 1.0])
    x = np.linspace(-eps, eps, 20001)
    d = E.fn1(D, x, eps)
    assert abs(np.trapezoid(d, x) - 1.0) < 1e-7
    h = E.fn1(H, x, eps)
    np.testing.assert_allclose(np.gradient(h, x)[5:-5], d[5:-5], atol=1e-6)
