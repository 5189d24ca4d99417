"""Pins of the implicit-path oracle (CPU only): golden CSR arrays and solution
values from the reference's tests, AMGCL built from the reference's own
headers (oracle/_ref), and operator consistency."""
import json
import os

import numpy as np
import pytest

import oracle
from helpers import desc

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CASES = [("SimplePoisson", [(1, 1), (1, 1)], False), ("SimplePoisson_Neum", [(2, 2), (2, 2)], True),
         ("SimplePoisson_Periodic", [(3, 3), (3, 3)], True)]


@pytest.mark.parametrize("name,bc,pin", CASES, ids=[c[0] for c in CASES])
def test_csr_golden(name, bc, pin):
    # CSRMatrixGeneratorTest.cpp:57-161: 4x4 cells on [0,4]^2, `1.0 == d2x(e) + d2y(e)`
    g = json.load(open(os.path.join(GOLD, "csr_simple_poisson.json")))["cases"][name]
    m = oracle.Mesh([5, 5], [(0., 4.), (0., 4.)])
    d = oracle.finalize(desc(2, [5, 5], 1, bc, ext=(1, 1)), m)
    ptr, col, val, rhs = oracle.csr_poisson(d, m, 1.0, pin)
    assert list(ptr) == g["ptr"]
    assert list(col) == g["col"]
    np.testing.assert_array_equal(val, g["val"])
    np.testing.assert_array_equal(rhs, g["rhs"])


def _csr_matvec(ptr, col, val, x):
    y = np.zeros(len(ptr) - 1)
    for r in range(len(ptr) - 1):
        y[r] = np.dot(val[ptr[r]:ptr[r + 1]], x[col[ptr[r]:ptr[r + 1]]])
    return y


def test_solution_golden_values():
    # AMGCLTest.cpp: x[0] == -0.375; HYPRE PCG+PFMG on the same system: x(1,1) = -1.125
    m = oracle.Mesh([5, 5], [(0., 4.), (0., 4.)])
    d = oracle.finalize(desc(2, [5, 5], 1, [(1, 1), (1, 1)], ext=(1, 1)), m)
    ptr, col, val, rhs = oracle.csr_poisson(d, m, 1.0, False)
    x, it, rel = oracle.pcg_mg([4, 4], [1.0, 1.0], [(1, 1), (1, 1)], rhs, tol=1e-13)
    assert x[0] == pytest.approx(-0.375, abs=1e-10)
    assert x[5] == pytest.approx(-1.125, abs=1e-10)
    assert rel <= 1e-13
    ref = oracle.amgcl_solve(ptr, col, val, rhs)
    if ref is not None:
        assert ref[0][0] == pytest.approx(-0.375, abs=1e-10)
        np.testing.assert_allclose(x, ref[0], rtol=0, atol=1e-10)


@pytest.mark.parametrize("bc", [[(1, 1), (1, 1)], [(3, 3), (1, 2)], [(2, 1), (1, 1)]])
def test_apply_matches_csr(bc):
    # the SPD operator (-d2 with folded ghosts) equals the assembled CSR matrix of `0 == d2x+d2y`
    n = 12
    m = oracle.Mesh([n + 1, n + 1], [(0., 1.), (0., 1.)])
    d = oracle.finalize(desc(2, [n + 1, n + 1], 1, bc, ext=(1, 1)), m)
    ptr, col, val, rhs = oracle.csr_poisson(d, m, 0.0, False)
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, n * n)
    y_csr = _csr_matvec(ptr, col, val, x)
    y = oracle.poisson_apply([n, n], [1.0 / n] * 2, bc, x)
    np.testing.assert_allclose(y, y_csr, rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("bc", [[(1, 1), (1, 1)], [(1, 2), (3, 3)]])
def test_pcg_mg_matches_amgcl(bc):
    n = 32
    m = oracle.Mesh([n + 1, n + 1])
    d = oracle.finalize(desc(2, [n + 1, n + 1], 1, bc, ext=(1, 1)), m)
    ptr, col, val, rhs0 = oracle.csr_poisson(d, m, 0.0, False)
    rng = np.random.default_rng(4)
    b = rng.normal(size=n * n)
    x, it, rel = oracle.pcg_mg([n, n], [1.0 / n] * 2, bc, b, tol=1e-12)
    assert rel <= 1e-12 and it < 60
    ref = oracle.amgcl_solve(ptr, col, val, b, tol=1e-13)
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    np.testing.assert_allclose(x, ref[0], rtol=0, atol=1e-9 * np.max(np.abs(ref[0])))
