"""The factored FTCS arithmetic (h a power of two, equal in every dimension) gives
the bits of D2SecondOrderCentered's operation order (D2SecondOrderCentered.hpp:160-171):
  reference  d2 = ((r-c)/h - (c-l)/h)/h per dimension, s = d2x + d2y (+ d2z), u + coef*s
  factored   D = (r-c) - (c-l),  S = Dx + Dy (+ Dz),  u + (coef/h^2)*S
Checked in numpy (IEEE float64 / float32, no FMA) on random data over many magnitudes,
including cancellation-heavy and subnormal inputs, against the reference order written
with divisions. Plausible slips (the wrong grouping of the sum, the textbook form
(r + l) - 2c) fail the same comparison; see test_factored_detects_regrouping."""
import numpy as np
import pytest


def reference(c, nb, h, coef, dt):
    h = dt(h)
    s = None
    for (l, r) in nb:
        d2 = ((r - c) / h - (c - l) / h) / h
        s = d2 if s is None else s + d2
    return c + dt(coef) * s


def factored(c, nb, h, coef, dt):
    coefh = dt(coef) * (dt(1.0 / h) * dt(1.0 / h))
    s = None
    for (l, r) in nb:
        d = (r - c) - (c - l)
        s = d if s is None else s + d
    return c + coefh * s


def _samples(rng, n, dt, scale):
    base = rng.uniform(-1, 1, n) * scale
    # neighbours close to the centre (cancellation) or unrelated
    near = lambda: base * (1 + rng.uniform(-1e-6, 1e-6, n))
    far = lambda: rng.uniform(-1, 1, n) * scale
    return base.astype(dt), [(near().astype(dt), far().astype(dt)), (far().astype(dt), near().astype(dt)),
                             (near().astype(dt), near().astype(dt))]


@pytest.mark.parametrize("dt,scales", [(np.float64, [1e-300, 1e-30, 1e-3, 1.0, 1e3, 1e30, 1e290]),
                                       (np.float32, [1e-30, 1e-3, 1.0, 1e3, 1e25])])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("k", [1, 6, 9, 10])
def test_factored_equals_reference(dt, scales, dim, k):
    rng = np.random.default_rng(100 * k + dim)
    h = 2.0 ** -k
    for coef_f in (0.2, 0.125, 1.0 / 6.0, 0.2499):
        coef = coef_f * h * h
        for sc in scales:
            c, nb = _samples(rng, 20000, dt, sc)
            with np.errstate(all="ignore"):
                ref = reference(c, nb[:dim], h, coef, dt)
                got = factored(c, nb[:dim], h, coef, dt)
            np.testing.assert_array_equal(got, ref)


def test_subnormal_inputs():
    rng = np.random.default_rng(7)
    tiny = np.float64(5e-324)
    c = (rng.integers(-1000, 1000, 5000) * tiny).astype(np.float64)
    nb = [((rng.integers(-1000, 1000, 5000) * tiny), (rng.integers(-1000, 1000, 5000) * tiny)) for _ in range(3)]
    h = 2.0 ** -9
    np.testing.assert_array_equal(factored(c, nb, h, 0.2 * h * h, np.float64),
                                  reference(c, nb, h, 0.2 * h * h, np.float64))


def test_factored_detects_regrouping():
    """The same check rejects a regrouped sum (Dx + (Dy + Dz)) and the textbook stencil
    form D = (r + l) - 2c, which differ from the reference only in rounding."""
    rng = np.random.default_rng(3)
    h = 2.0 ** -9
    coef = 0.2 * h * h
    c, nb = _samples(rng, 20000, np.float64, 1.0)
    ref = reference(c, nb, h, coef, np.float64)
    d = [(r - c) - (c - l) for (l, r) in nb]
    regrouped = c + coef / h / h * (d[0] + (d[1] + d[2]))
    textbook = c + coef / h / h * sum((r + l) - 2 * c for (l, r) in nb)
    assert not np.array_equal(regrouped, ref)
    assert not np.array_equal(textbook, ref)
