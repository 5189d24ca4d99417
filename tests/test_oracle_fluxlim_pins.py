"""Pins of the flux-limiter restatement in oracle/expr.py (CPU only).

OpFlow has no unit test for D1FluxLimiter / FluxLimiterKernels, so each part is
pinned by values the mathematics fixes, derived here independently:
  * the limiter functions psi(r): textbook values at sample ratios (Roe's
    superbee, van Leer's harmonic limiter, van Albada, minmod, MUSCL), psi(1) = 1
    (second order on smooth data), psi(r) = 0 for r <= 0 where the limiter clips,
    the Sweby TVD region 0 <= psi <= min(2r, 2), and the symmetry
    psi(r)/r = psi(1/r) of the symmetric limiters;
  * the face interpolation: exact face values of linear data on a NON-uniform
    mesh for every kernel, upwind and downwind, Cen2Cor and Cor2Cen (fixes the
    point offsets, the centre coordinates and the dx that multiplies K), the
    kappa schemes' exactness on quadratic data where the scheme is third order
    (QUICK on a uniform mesh), and the choice of branch by the sign of u;
  * conditional / min / max / comparisons: numpy's own functions.
The SPL / GPL kernels read their std::ratio parameters as integer divisions in
the reference (FluxLimiterKernels.hpp:76-79, 128-130), so kappa and alpha of
ratios below one are 0; their pins use those values.
"""
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import expr as E

PAD = 4
TVD = [E.LIM_MINMOD, E.LIM_SUPERBEE, E.LIM_MUSCL, E.LIM_HARMONIC, E.LIM_UMIST, E.LIM_SPL13, E.LIM_SPLMAX12,
       E.LIM_SPLMAX13, E.LIM_KOREN, E.LIM_BSOU]
RATIO = TVD + [E.LIM_VANALBADA, E.LIM_SMART]
ALL = list(range(17))


def test_limiter_textbook_values():
    r = np.array([0.5, 2.0, 3.0])
    ex = {
        E.LIM_MINMOD: [0.5, 1.0, 1.0],            # max(0, min(r, 1))
        E.LIM_SUPERBEE: [1.0, 2.0, 2.0],          # max(0, min(2r, 1), min(r, 2))
        E.LIM_MUSCL: [0.75, 1.5, 2.0],            # max(0, min(2r, (1 + r)/2, 2))
        E.LIM_HARMONIC: [2 / 3, 4 / 3, 1.5],      # van Leer: 2r / (1 + r) for r > 0
        E.LIM_VANALBADA: [0.6, 1.2, 1.2],         # (r^2 + r) / (r^2 + 1)
        E.LIM_KOREN: [0.75, 1.5, 2.0],            # kappa read as 0: max(0, min(2r, (1 + r)/2, 2))
        E.LIM_SMART: [0.75, 1.5, 2.0],            # min(2r, (1 + r)/2, 4)
        E.LIM_BSOU: [1.0, 1.0, 1.0],              # kappa = -1: min(2r, 1, 1)
    }
    for k, v in ex.items():
        np.testing.assert_allclose(E.limiter(k, r), v, rtol=1e-15, err_msg=str(k))


@pytest.mark.parametrize("k", RATIO)
def test_limiter_second_order_at_one(k):
    assert E.limiter(k, np.array([1.0]))[0] == 1.0


@pytest.mark.parametrize("k", TVD)
def test_limiter_tvd_region(k):
    r = np.concatenate([-np.logspace(-3, 3, 50), np.logspace(-3, 3, 200)])
    p = E.limiter(k, r)
    neg = r <= 0
    assert np.all(p[neg] == 0.0)
    assert np.all(p[~neg] >= 0.0) and np.all(p[~neg] <= np.minimum(2 * r[~neg], 2.0) * (1 + 1e-15))


@pytest.mark.parametrize("k", [E.LIM_MINMOD, E.LIM_SUPERBEE, E.LIM_MUSCL, E.LIM_HARMONIC, E.LIM_VANALBADA])
def test_limiter_symmetry(k):
    r = np.logspace(-2, 2, 101)
    np.testing.assert_allclose(E.limiter(k, r) / r, E.limiter(k, 1 / r), rtol=1e-14)


class Mesh1:
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


def _interp(mesh, vel, vals, loc, kernel, lo, hi):
    """evaluate() of FLUXLIM(u, e) on 1-D fields stored from index -PAD."""
    du = SimpleNamespace(offset=[-PAD], extent=[len(vel)], loc=[1 - loc])
    de = SimpleNamespace(offset=[-PAD], extent=[len(vals)], loc=[loc])
    nodes = [(E.FIELD, 0, 0, 0, 0.0), (E.FIELD, 1, 0, 0, 0.0), (E.FLUXLIM, 0, 1, 0, float(kernel))]
    return E.evaluate(nodes, 2, [(np.asarray(vel, float), du), (np.asarray(vals, float), de)], mesh, [lo], [hi])


@pytest.mark.parametrize("kernel", ALL)
@pytest.mark.parametrize("loc", [E.CENTER, E.CORNER], ids=["cen2cor", "cor2cen"])
@pytest.mark.parametrize("sign", [1.0, -1.0], ids=["upwind", "downwind"])
def test_linear_data_exact_on_nonuniform_mesh(kernel, loc, sign):
    n = 24
    rng = np.random.default_rng(5)
    nodes = np.concatenate([[0.0], np.cumsum(rng.uniform(0.6, 1.4, n))]) / n
    mesh = Mesh1(nodes)
    idx = np.arange(-PAD, n + PAD)
    pts = mesh.x(0, idx) + (0.5 * mesh.dx(0, idx) if loc == E.CENTER else 0.0)
    vals = 0.7 + 2.5 * pts
    vel = np.full(len(idx), sign)
    got = _interp(mesh, vel, vals, loc, kernel, 3, n - 3)
    i = np.arange(3, n - 3)
    target = mesh.x(0, i) if loc == E.CENTER else mesh.x(0, i) + 0.5 * mesh.dx(0, i)  # faces / centres
    np.testing.assert_allclose(got, 0.7 + 2.5 * target, rtol=0, atol=1e-12)


def test_quick_exact_on_quadratic_data():
    """QUICK (kappa = 1/2) is the quadratic through the three upwind point values, so on
    a uniform mesh it reproduces a quadratic's face values exactly, upwind and downwind;
    Fromm (kappa = 0) and the central scheme (kappa = 1) do not, which pins the split of
    the two slopes between (1 + kappa)/2 and (1 - kappa)/2."""
    n = 32
    mesh = Mesh1(np.linspace(0.0, 1.0, n + 1))
    idx = np.arange(-PAD, n + PAD)
    c = mesh.x(0, idx) + 0.5 * mesh.dx(0, idx)
    vals = 0.3 - c + 2.0 * c * c
    faces = mesh.x(0, np.arange(3, n - 3))
    exact = 0.3 - faces + 2.0 * faces * faces
    for sign in (1.0, -1.0):
        got = _interp(mesh, np.full(len(idx), sign), vals, E.CENTER, E.LIM_QUICK, 3, n - 3)
        np.testing.assert_allclose(got, exact, rtol=0, atol=1e-14)
        for k in (E.LIM_FROMM, E.LIM_CDS):
            other = _interp(mesh, np.full(len(idx), sign), vals, E.CENTER, k, 3, n - 3)
            assert np.max(np.abs(other - exact)) > 1e-5


def test_branch_follows_velocity_sign():
    """u > 0 takes the upwind stencil (points i-2, i-1, i for Cen2Cor), u <= 0 the downwind one
    (i-1, i, i+1): a jump placed so only one stencil sees it tells them apart."""
    n = 16
    mesh = Mesh1(np.linspace(0.0, 1.0, n + 1))
    idx = np.arange(-PAD, n + PAD)
    vals = np.where(idx >= 8, 1.0, 0.0)
    vel = np.where(np.arange(len(idx)) % 2 == 0, 1.0, -1.0)
    got = _interp(mesh, vel, vals, E.CENTER, E.LIM_MINMOD, 2, n - 2)
    i = np.arange(2, n - 2)
    up = i - 1 >= 8  # upwind face value is the left cell's (minmod: slope limited to 0 at the jump)
    dn = i >= 8
    want = np.where(vel[i + PAD] > 0, up.astype(float), dn.astype(float))
    np.testing.assert_array_equal(got, want)


def test_conditional_min_max_compare():
    desc = SimpleNamespace(offset=[0], extent=[50], loc=[E.CENTER])
    mesh = Mesh1(np.linspace(0, 1, 51))
    rng = np.random.default_rng(1)
    a, b = rng.normal(size=50), rng.normal(size=50)
    b[::7] = a[::7]  # ties
    fields = [(a, desc), (b, desc)]
    base = [(E.FIELD, 0, 0, 0, 0.0), (E.FIELD, 1, 0, 0, 0.0)]
    for kind, ref in ((E.MIN, np.minimum(a, b)), (E.MAX, np.maximum(a, b)), (E.GT, (a > b) * 1.0),
                      (E.LT, (a < b) * 1.0), (E.GE, (a >= b) * 1.0), (E.LE, (a <= b) * 1.0)):
        got = E.evaluate(base + [(kind, 0, 1, 0, 0.0)], 2, fields, mesh, [0], [50])
        np.testing.assert_array_equal(got, ref)
    nodes = base + [(E.GT, 0, 1, 0, 0.0), (E.COND, 2, 0, 1, 0.0)]  # a > b ? a : b
    np.testing.assert_array_equal(E.evaluate(nodes, 3, fields, mesh, [0], [50]), np.where(a > b, a, b))
    nodes = base + [(E.ABS, 0, 0, 0, 0.0)]
    np.testing.assert_array_equal(E.evaluate(nodes, 2, fields, mesh, [0], [50]), np.abs(a))
