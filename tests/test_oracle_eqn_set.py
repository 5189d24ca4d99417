"""Pins of the oracle's equation-set assembly (CPU only): oracle.expr.linear_system
against the reference's golden CSR arrays (CSRMatrixGeneratorTest.cpp:57-161), the
block structure its two-equation tests assert (:163-250), and closed forms of
coupling blocks."""
import json
import os

import numpy as np
import pytest

import oracle
from oracle import expr as X
from helpers import desc

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CASES = [("SimplePoisson", [(1, 1), (1, 1)], False), ("SimplePoisson_Neum", [(2, 2), (2, 2)], True),
         ("SimplePoisson_Periodic", [(3, 3), (3, 3)], True)]


def _dense(g):
    n = len(g["ptr"]) - 1
    A = np.zeros((n, n))
    for r in range(n):
        for k in range(g["ptr"][r], g["ptr"][r + 1]):
            A[r, g["col"][k]] = g["val"][k]
    return A, np.asarray(g["rhs"])


def _poisson_nodes(u, c=1.0):
    # c == d2x(e_u) + d2y(e_u)
    return [(X.CONST, 0, 0, 0, c), (X.UNKNOWN, u, 0, 0, 0.0), (X.D2C, 1, 0, 0, 0.0), (X.D2C, 1, 0, 1, 0.0),
            (X.ADD, 2, 3, 0, 0.0)], 0, 4


def _setup(bc, n=4):
    m = oracle.Mesh([n + 1, n + 1], [(0., float(n)), (0., float(n))])
    d = oracle.finalize(desc(2, [n + 1, n + 1], 1, bc, ext=(1, 1)), m)
    arr = np.zeros((d.extent[1], d.extent[0]))
    return m, d, arr


@pytest.mark.parametrize("name,bc,pin", CASES, ids=[c[0] for c in CASES])
def test_linear_system_matches_golden_csr(name, bc, pin):
    g = json.load(open(os.path.join(GOLD, "csr_simple_poisson.json")))["cases"][name]
    m, d, arr = _setup(bc)
    nodes, lr, rr = _poisson_nodes(0)
    A, rhs = X.linear_system([(nodes, lr, rr, [])], [(arr, d)], m, lambda i, a: oracle.update_padding(d, m, a),
                             [pin])
    Ag, rg = _dense(g)
    np.testing.assert_array_equal(A, Ag)
    np.testing.assert_array_equal(rhs, rg)


@pytest.mark.parametrize("name,bc,pin", CASES[:2], ids=[c[0] for c in CASES[:2]])
def test_two_equation_blocks(name, bc, pin):
    # SimplePoisson_2Eqn / _Neum_2Eqn: the second block is the first shifted by 16 columns
    g = json.load(open(os.path.join(GOLD, "csr_simple_poisson.json")))["cases"][name]
    m, d, arr = _setup(bc)
    e0 = _poisson_nodes(0)
    e1 = _poisson_nodes(1)
    A, rhs = X.linear_system([(*e0, []), (*e1, [])], [(arr, d), (arr, d)], m,
                             lambda i, a: oracle.update_padding(d, m, a), [pin, pin])
    Ag, rg = _dense(g)
    np.testing.assert_array_equal(A[:16, :16], Ag)
    np.testing.assert_array_equal(A[16:, 16:], Ag)
    assert not A[:16, 16:].any() and not A[16:, :16].any()
    np.testing.assert_array_equal(rhs, np.concatenate([rg, rg]))


def test_coupling_blocks_closed_forms():
    # eq0: 1 == lap(e0) - 3 e1      -> block (0,1) = +3 I
    # eq1: 0.5 == lap(e1) + D1Up_x(e0) (Dirichlet 0: ghost = -mirror)
    #   -> block (1,0) row r: -(e0[r+1] - e0[r]) / h, the last cell of a row -(-2 e0[r]) / h
    m, d, arr = _setup([(1, 1), (1, 1)])
    n0 = [(X.CONST, 0, 0, 0, 1.0), (X.UNKNOWN, 0, 0, 0, 0.0), (X.D2C, 1, 0, 0, 0.0), (X.D2C, 1, 0, 1, 0.0),
          (X.ADD, 2, 3, 0, 0.0), (X.UNKNOWN, 1, 0, 0, 0.0), (X.CONST, 0, 0, 0, 3.0), (X.MUL, 6, 5, 0, 0.0),
          (X.SUB, 4, 7, 0, 0.0)]
    n1 = [(X.CONST, 0, 0, 0, 0.5), (X.UNKNOWN, 1, 0, 0, 0.0), (X.D2C, 1, 0, 0, 0.0), (X.D2C, 1, 0, 1, 0.0),
          (X.ADD, 2, 3, 0, 0.0), (X.UNKNOWN, 0, 0, 0, 0.0), (X.D1UP, 5, 0, 0, 0.0), (X.ADD, 4, 6, 0, 0.0)]
    A, rhs = X.linear_system([(n0, 0, 8, []), (n1, 0, 7, [])], [(arr, d), (arr, d)], m,
                             lambda i, a: oracle.update_padding(d, m, a))
    np.testing.assert_array_equal(A[:16, 16:], 3.0 * np.eye(16))
    B = np.zeros((16, 16))
    for r in range(16):
        i = r % 4
        if i < 3:
            B[r, r], B[r, r + 1] = 1.0, -1.0
        else:
            B[r, r] = 2.0
    np.testing.assert_array_equal(A[16:, :16], B)
    # the diagonal blocks are the golden Dirichlet Laplacian rows
    g = json.load(open(os.path.join(GOLD, "csr_simple_poisson.json")))["cases"]["SimplePoisson"]
    Ag, rg = _dense(g)
    np.testing.assert_array_equal(A[:16, :16], Ag)
    np.testing.assert_array_equal(A[16:, 16:], Ag)
    np.testing.assert_array_equal(rhs, np.concatenate([rg, 0.5 * rg]))


def test_linear_system_is_the_evaluated_affine_map():
    # F(u) = lhs - rhs over random interiors equals A u - rhs, inhomogeneous boundaries
    m = oracle.Mesh([7, 6], [(0., 1.), (0., 2.)])
    bcv = [(0.7, -0.3), (1.1, 0.4)]
    d = oracle.finalize(desc(2, [7, 6], 1, [(1, 2), (2, 1)], bcv=bcv, ext=(1, 1)), m)
    arr = np.zeros((d.extent[1], d.extent[0]))
    n0 = [(X.CONST, 0, 0, 0, 1.0), (X.UNKNOWN, 0, 0, 0, 0.0), (X.D2C, 1, 0, 0, 0.0), (X.D2C, 1, 0, 1, 0.0),
          (X.ADD, 2, 3, 0, 0.0), (X.UNKNOWN, 1, 0, 0, 0.0), (X.D1DOWN, 5, 0, 1, 0.0), (X.SUB, 4, 6, 0, 0.0)]
    fill = lambda i, a: oracle.update_padding(d, m, a)  # noqa: E731
    A, rhs = X.linear_system([(n0, 0, 7, [])], [(arr, d), (arr, d)], m, fill)
    rng = np.random.default_rng(5)
    N = 30
    u = [rng.standard_normal(N), rng.standard_normal(N)]
    sub = [(X.FIELD, a, 0, 0, 0.0) if k == X.UNKNOWN else (k, a, b, dd, v) for (k, a, b, dd, v) in n0]
    sub.append((X.SUB, 0, 7, 0, 0.0))
    sl = (slice(1, 6), slice(1, 7))
    fields = []
    for j in range(2):
        a = np.zeros_like(arr)
        a[sl] = u[j].reshape(5, 6)
        fields.append((fill(j, a), d))
    Fu = X.evaluate(sub, len(sub) - 1, fields, m, [0, 0], [6, 5]).reshape(-1)
    np.testing.assert_allclose(A[:N] @ np.concatenate(u) - rhs[:N], Fu, rtol=0, atol=1e-12)
