"""Pins of the operator restatements in oracle/expr.py (CPU only).

OpFlow has no unit tests for these operators, so each is pinned by what the
mathematics fixes, derived here independently of the restated formulas:
  * first/second differences and linear interpolation (D1FirstOrderCentered,
    D1FirstOrderBiasedUpwind/Downwind, D2SecondOrderCentered, D1Linear): exact
    values on linear / quadratic data on a NON-uniform mesh, which fixes the
    direction, the offsets and every mesh-spacing term;
  * WENO53 (D1WENO53Upwind/Downwind): the candidate derivatives and the
    smoothness indicators against their definitions (Jiang & Shu 1996: the
    quadratic with given cell averages, its value at the cell's right face and
    sum_l int (d^l p / dx^l)^2 over the cell), the linear weights against the
    quartic reconstruction from five cell averages, the points each candidate
    uses against polynomial interpolation (np.polyfit) through the evaluator's
    window, and fifth-order convergence on smooth data.
A dropped term, a wrong sign or index, a swapped weight or a window shifted by
one fails at least one of these.
"""
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import expr as E

PAD = 4


class Mesh1:
    """1-D mesh with arbitrary node coordinates (x(i) = node i, dx(i) = x(i+1) - x(i))."""

    def __init__(self, nodes):
        self.dim = 1
        lo = nodes[0] - (nodes[1] - nodes[0]) * np.arange(PAD, 0, -1)
        hi = nodes[-1] + (nodes[-1] - nodes[-2]) * np.arange(1, PAD + 1)
        self.xs = np.concatenate([lo, nodes, hi])

    def x(self, d, i):
        return self.xs[np.asarray(i) + PAD]

    def dx(self, d, i):
        i = np.asarray(i) + PAD
        return self.xs[i + 1] - self.xs[i]


def _nonuniform(n, seed):
    rng = np.random.default_rng(seed)
    return np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 1.5, n - 1))]) / n


def _eval1(mesh, values, loc, kind, lo, hi, offset=-PAD):
    """evaluate() of one operator node on a 1-D field holding `values` from index `offset`."""
    desc = SimpleNamespace(offset=[offset], extent=[len(values)], loc=[loc])
    nodes = [(E.FIELD, 0, 0, 0, 0.0), (kind, 0, 0, 0, 0.0)]
    return E.evaluate(nodes, 1, [(np.asarray(values, float), desc)], mesh, [lo], [hi])


def _sample(mesh, n, loc, f):
    idx = np.arange(-PAD, n + PAD - 1)
    pts = mesh.x(0, idx) + (0.5 * mesh.dx(0, idx) if loc == E.CENTER else 0.0)
    return f(pts), idx


# ------------------------------------------------------------------ differences and interpolation
A, B, G = 0.7, -1.3, 2.1  # u = A + B x + G x^2


@pytest.mark.parametrize("loc", [E.CENTER, E.CORNER], ids=["center", "corner"])
def test_biased_differences_exact_on_quadratics(loc):
    n = 24
    m = Mesh1(_nonuniform(n, 1))
    u, idx = _sample(m, n, loc, lambda x: A + B * x + G * x * x)
    p = (m.x(0, idx) + (0.5 * m.dx(0, idx) if loc == E.CENTER else 0.0))  # sample points by index
    i = np.arange(2, n - 2)
    at = lambda k: p[i + k + PAD]
    up = _eval1(m, u, loc, E.D1UP, 2, n - 2)
    down = _eval1(m, u, loc, E.D1DOWN, 2, n - 2)
    # a forward / backward difference of a quadratic divided by the true distance
    np.testing.assert_allclose(up, B + G * (at(0) + at(1)), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(down, B + G * (at(-1) + at(0)), rtol=1e-12, atol=1e-12)


def test_centered_first_difference_staggered():
    n = 24
    m = Mesh1(_nonuniform(n, 2))
    i = np.arange(2, n - 2)
    # centre -> corner: (u_i - u_{i-1}) over the distance of the two centres
    u, idx = _sample(m, n, E.CENTER, lambda x: A + B * x + G * x * x)
    c = m.x(0, idx) + 0.5 * m.dx(0, idx)
    got = _eval1(m, u, E.CENTER, E.D1C, 2, n - 2)
    np.testing.assert_allclose(got, B + G * (c[i - 1 + PAD] + c[i + PAD]), rtol=1e-12, atol=1e-12)
    # corner -> centre: (u_{i+1} - u_i) / dx_i
    u, idx = _sample(m, n, E.CORNER, lambda x: A + B * x + G * x * x)
    x = m.x(0, idx)
    got = _eval1(m, u, E.CORNER, E.D1C, 2, n - 2)
    np.testing.assert_allclose(got, B + G * (x[i + PAD] + x[i + 1 + PAD]), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("loc", [E.CENTER, E.CORNER], ids=["center", "corner"])
def test_second_difference_exact_on_quadratics(loc):
    n = 24
    m = Mesh1(_nonuniform(n, 3))
    u, _ = _sample(m, n, loc, lambda x: A + B * x + G * x * x)
    got = _eval1(m, u, loc, E.D2C, 2, n - 2)
    np.testing.assert_allclose(got, 2 * G, rtol=1e-10)


def test_linear_interpolation_exact_on_linear_data():
    n = 24
    m = Mesh1(_nonuniform(n, 4))
    i = np.arange(2, n - 2)
    # centre -> node i from the centres of cells i-1 and i (unequal widths: the weights matter)
    u, _ = _sample(m, n, E.CENTER, lambda x: A + B * x)
    got = _eval1(m, u, E.CENTER, E.INTP_C2N, 2, n - 2)
    np.testing.assert_allclose(got, A + B * m.x(0, i), rtol=1e-12)
    # node -> centre of cell i: the midpoint of nodes i and i+1
    u, _ = _sample(m, n, E.CORNER, lambda x: A + B * x)
    got = _eval1(m, u, E.CORNER, E.INTP_N2C, 2, n - 2)
    np.testing.assert_allclose(got, A + B * (m.x(0, i) + 0.5 * m.dx(0, i)), rtol=1e-12)
    # a linear interpolation is not exact for a quadratic: the test data can tell
    u, _ = _sample(m, n, E.CENTER, lambda x: x * x)
    got = _eval1(m, u, E.CENTER, E.INTP_C2N, 2, n - 2)
    assert np.max(np.abs(got - m.x(0, i) ** 2)) > 1e-6


# ------------------------------------------------------------------ WENO53
def _cell_avg_quadratic(avgs, centres):
    """Coefficients (c0, c1, c2) of p(x) = c0 + c1 x + c2 x^2 whose averages over the unit
    cells centred at `centres` are `avgs` (the cell-average form of the reconstruction)."""
    M = np.array([[1.0, c, c * c + 1.0 / 12.0] for c in centres])
    return np.linalg.solve(M, np.asarray(avgs, float))


def _indicator(coef):
    """sum_{l=1,2} int_{-1/2}^{1/2} (p^(l))^2 dx for the quadratic p (unit cell)."""
    c0, c1, c2 = coef
    # p' = c1 + 2 c2 x: int (p')^2 = c1^2 + (4 c2^2)/12 ; p'' = 2 c2: int (p'')^2 = 4 c2^2
    return c1 * c1 + 4.0 * c2 * c2 / 12.0 + 4.0 * c2 * c2


def test_weno_candidates_and_indicators_match_their_definitions():
    rng = np.random.default_rng(11)
    h = 1.0 / 32
    for _ in range(200):
        d = rng.normal(size=5)  # d1..d5: cell averages of cells -2..2 (the derivative values)
        # six upwind-window values whose forward differences / h are d5, d4, d3, d2, d1 (index order)
        phi = np.concatenate([[0.0], np.cumsum(h * d[::-1])])
        (c1, c2, c3), (s1, s2, s3), _ = E._weno_parts(list(phi), h, True)
        stencils = [(d[0:3], [-2, -1, 0]), (d[1:4], [-1, 0, 1]), (d[2:5], [0, 1, 2])]
        for (avg, cen), cand, ind in zip(stencils, (c1, c2, c3), (s1, s2, s3)):
            coef = _cell_avg_quadratic(avg, cen)
            assert cand == pytest.approx(coef[0] + coef[1] * 0.5 + coef[2] * 0.25, rel=1e-10, abs=1e-10)
            assert ind == pytest.approx(_indicator(coef), rel=1e-10, abs=1e-10)


def test_weno_linear_weights_give_the_fifth_order_reconstruction():
    rng = np.random.default_rng(12)
    cen = [-2, -1, 0, 1, 2]
    M = np.array([[((c + 0.5) ** (k + 1) - (c - 0.5) ** (k + 1)) / (k + 1) for k in range(5)] for c in cen])
    for _ in range(50):
        d = rng.normal(size=5)
        quartic = np.linalg.solve(M, d)  # polynomial with the five cell averages
        target = sum(quartic[k] * 0.5 ** k for k in range(5))
        cands = [_cell_avg_quadratic(d[i:i + 3], cen[i:i + 3]) for i in range(3)]
        vals = [c[0] + c[1] * 0.5 + c[2] * 0.25 for c in cands]
        assert 0.1 * vals[0] + 0.6 * vals[1] + 0.3 * vals[2] == pytest.approx(target, rel=1e-10, abs=1e-10)


def test_weno_uses_the_linear_weights_when_the_indicators_agree():
    # differences (d1..d5) = (1/4, 0, 1, 0, 1/4): all three indicators equal 13/3 exactly
    h = 1.0 / 32
    d = np.array([0.25, 0.0, 1.0, 0.0, 0.25])
    phi = np.concatenate([[0.0], np.cumsum(h * d[::-1])])
    cands, inds, w = E._weno_parts(list(phi), h, True)
    assert inds[0] == pytest.approx(13 / 3, rel=1e-14) and inds[1] == pytest.approx(13 / 3, rel=1e-14)
    assert inds[2] == pytest.approx(13 / 3, rel=1e-14)
    np.testing.assert_allclose(w, [0.1, 0.6, 0.3], rtol=1e-14)
    assert E._weno(list(phi), h, True) == pytest.approx(0.1 * cands[0] + 0.6 * cands[1] + 0.3 * cands[2], rel=1e-14)


def _weno_eval(phi, kind, lo, hi, h, offset):
    m = SimpleNamespace(dim=1, x=lambda d, i: np.asarray(i) * h, dx=lambda d, i: np.full(np.shape(i), h))
    desc = SimpleNamespace(offset=[offset], extent=[len(phi)], loc=[E.CENTER])
    nodes = [(E.FIELD, 0, 0, 0, 0.0), (kind, 0, 0, 0, 0.0)]
    return E.evaluate(nodes, 1, [(np.asarray(phi, float), desc)], m, [lo], [hi])


@pytest.mark.parametrize("kind", [E.WENO_UP, E.WENO_DOWN], ids=["upwind", "downwind"])
def test_weno_window_and_candidates_by_isolation(kind):
    """Jumps that put two of the three candidates across a discontinuity leave the
    third: the result is then the derivative at x_i of the cubic through that
    candidate's four points (np.polyfit), which fixes the evaluator's window."""
    h = 1.0 / 32
    i = 8
    rng = np.random.default_rng(13)
    # candidate point sets relative to i (upwind window i-2..i+3, downwind i-3..i+2)
    if kind == E.WENO_UP:
        sets = {(-1, 0): [0, 1, 2, 3], (1, 2): [-2, -1, 0, 1], (-2, -1, 2, 3): [-1, 0, 1, 2]}
    else:
        sets = {(0, 1): [-3, -2, -1, 0], (-2, -1): [-1, 0, 1, 2], (-3, -2, 1, 2): [-2, -1, 0, 1]}
    for jumps, pts in sets.items():
        pts = sorted(pts)
        for _ in range(5):
            phi = rng.uniform(-1, 1, 20)
            for a in range(0, len(jumps), 2):  # a jump of 1e4 between i + jumps[a] and i + jumps[a+1]
                phi[i + jumps[a + 1]:] += 1e4
            got = _weno_eval(phi, kind, i, i + 1, h, 0)[0]
            xs = np.array([(i + p) * h for p in pts])
            poly = np.polyfit(xs - i * h, phi[[i + p for p in pts]], 3)
            want = poly[2]  # derivative at x_i of the interpolating cubic
            assert got == pytest.approx(want, rel=1e-6, abs=1e-6 * abs(want) + 1e-6), (jumps, pts)


@pytest.mark.parametrize("kind", [E.WENO_UP, E.WENO_DOWN], ids=["upwind", "downwind"])
def test_weno_exact_on_cubics_and_fifth_order(kind):
    rng = np.random.default_rng(14)
    for _ in range(5):
        c = rng.normal(size=4)
        h = 1.0 / 32
        x = np.arange(-PAD, 40) * h
        phi = c[0] + c[1] * x + c[2] * x ** 2 + c[3] * x ** 3
        got = _weno_eval(phi, kind, 4, 30, h, -PAD)
        xi = np.arange(4, 30) * h
        np.testing.assert_allclose(got, c[1] + 2 * c[2] * xi + 3 * c[3] * xi ** 2, rtol=1e-9, atol=1e-9)
    # smooth data: the error falls ~32x per halving of h (5th order); a swapped weight or a
    # dropped indicator term lowers the order
    errs = []
    for n in (256, 512, 1024):
        h = 1.0 / n
        x = np.arange(-PAD, n + PAD) * h
        phi = np.sin(2 * np.pi * x) + 0.5 * np.cos(6 * np.pi * x)
        k = n // 8  # x = 1/8 at every resolution, away from the critical points of the function
        got = _weno_eval(phi, kind, k, k + 1, h, -PAD)[0]
        xk = k * h
        errs.append(abs(got - (2 * np.pi * np.cos(2 * np.pi * xk) - 3 * np.pi * np.sin(6 * np.pi * xk))))
    assert errs[0] / errs[1] > 28 and errs[1] / errs[2] > 28, errs
