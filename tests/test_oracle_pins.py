"""Pins of the oracle (CPU only): golden values from the reference's own
tests, closed forms and brute force. A plausible mistake in the restatement
(sign, index, dropped term, transposed operand) fails one of these."""
import json
import os

import numpy as np
import pytest

import oracle
from opflow_amd import _lib as L
from helpers import desc

GOLD = os.path.join(os.path.dirname(__file__), "golden")
D, N, P = L.OPF_LOC_CORNER if hasattr(L, "OPF_LOC_CORNER") else 0, 2, 3
DIRC, NEUM, PER, SYMM, ASYMM, UNDEF = 1, 2, 3, 5, 6, 0


def orc_desc(dim, nodes, loc, bc, bcv=None, ext=(1, 1), padding=0):
    d = desc(dim, nodes, loc, bc, bcv, ext, padding)
    return oracle.finalize(d, oracle.Mesh(nodes))


@pytest.mark.parametrize("case", json.load(open(os.path.join(GOLD, "field_ranges.json")))["cases"],
                         ids=lambda c: c["name"])
def test_ranges_golden(case):
    # CartesianFieldTest.cpp:33-146: 1-D mesh with 11 nodes
    m = oracle.Mesh([11])
    d = desc(1, [11], case["loc"], [case["bc"]], ext=tuple(case["ext"]))
    o = oracle.finalize(d, m)
    for key, (a, b) in [("accessible", ("access_start", "access_end")), ("local", ("local_start", "local_end")),
                        ("logical", ("logical_start", "logical_end")),
                        ("assignable", ("assign_start", "assign_end"))]:
        exp = case[key]
        assert getattr(o, a)[0] == 0 + exp[0], key
        assert getattr(o, b)[0] == 11 + exp[1], key
        # the product's host-side finalize agrees
        assert getattr(d, a)[0] == getattr(o, a)[0]
        assert getattr(d, b)[0] == getattr(o, b)[0]


def test_dirc_ghost_linear_exact():
    # DircBCTest.cpp:27-56: u = x extrapolates exactly into the ghosts, corner and centre
    m = oracle.Mesh([11])
    for loc in (0, 1):
        d = orc_desc(1, [11], loc, [(DIRC, DIRC)], [(0.0, 1.0)], ext=(1, 1))
        u = np.zeros(d.extent[0])
        for i in range(d.local_start[0], d.local_end[0]):
            u[i - d.offset[0]] = m.x(0, i) + (0.5 * m.dx(0, i) if loc == 1 else 0.0)
        u = oracle.update_padding(d, m, u)
        for i in range(d.logical_start[0], d.logical_end[0]):
            exp = m.x(0, i) + (m.dx(0, i) / 2. if loc == 1 else 0.0)
            assert u[i - d.offset[0]] == pytest.approx(exp, rel=1e-14, abs=1e-15), (loc, i)


def test_center_periodic_ghost_values():
    # CartesianFieldTest.cpp:161-176: u = i1*10 + i0 on a 10x10 centre periodic field
    m = oracle.Mesh([11, 11])
    d = orc_desc(2, [11, 11], 1, [(PER, PER), (PER, PER)], ext=(1, 1))
    u = np.zeros((d.extent[1], d.extent[0]))
    for j in range(10):
        for i in range(10):
            u[j - d.offset[1], i - d.offset[0]] = j * 10 + i
    u = oracle.update_padding(d, m, u)
    for j in range(-1, 11):
        for i in range(-1, 11):
            assert u[j - d.offset[1], i - d.offset[0]] == (j + 10) % 10 * 10 + (i + 10) % 10


def test_reduce_known_sums():
    # RangeReduceTest.cpp:37-60: 1 + 10 + ... + 1e9 and a field of ones
    m = oracle.Mesh([11])
    d = orc_desc(1, [11], 1, [(DIRC, DIRC)], ext=(0, 0))
    v = np.array([10.0 ** i for i in range(10)])
    u = np.zeros(d.extent[0])
    u[0 - d.offset[0]:10 - d.offset[0]] = v
    assert oracle.reduce(d, u, "sum", [0], [10]) == 1111111111.0
    m2 = oracle.Mesh([11, 11])
    d2 = orc_desc(2, [11, 11], 1, [(DIRC, DIRC), (DIRC, DIRC)], ext=(0, 0))
    ones = np.ones((d2.extent[1], d2.extent[0]))
    assert oracle.reduce(d2, ones, "sum", [0, 0], [10, 10]) == 100.0


def _dirichlet_eigen_decay(n, steps, m_, k_):
    h = 1.0 / n
    dt = 0.2 * h * h
    lam = 1 - 4 * (dt / h ** 2) * (np.sin(m_ * np.pi * h / 2) ** 2 + np.sin(k_ * np.pi * h / 2) ** 2)
    return lam ** steps


def test_ftcs2d_eigenmode_closed_form():
    # sin(m pi x) sin(k pi y) at cell centres with Dirichlet-0 ghosts (ghost = -mirror) is an exact
    # discrete eigenmode of the 5-point D2 operator: u^N = lambda^N u^0.
    n, steps, mm, kk = 32, 50, 3, 2
    mesh = oracle.Mesh([n + 1, n + 1])
    d = orc_desc(2, [n + 1, n + 1], 1, [(DIRC, DIRC), (DIRC, DIRC)], ext=(1, 1))
    xc = (np.arange(n) + 0.5) / n
    u0 = np.zeros((d.extent[1], d.extent[0]))
    u0[1:-1, 1:-1] = np.outer(np.sin(kk * np.pi * xc), np.sin(mm * np.pi * xc))
    u0 = oracle.update_padding(d, mesh, u0)
    h = 1.0 / n
    u = oracle.ftcs(d, mesh, u0, 0.2 * h * h, steps)
    exp = _dirichlet_eigen_decay(n, steps, mm, kk) * u0[1:-1, 1:-1]
    assert np.max(np.abs(u[1:-1, 1:-1] - exp)) <= 1e-13 * np.max(np.abs(exp))


def test_ftcs3d_periodic_fourier_closed_form():
    # cos(2 pi (a x + b y + c z)) on a periodic grid: lambda = 1 - 4 r sum sin^2(pi k_d h)
    n, steps = 16, 20
    a, b, c = 1, 2, 3
    mesh = oracle.Mesh([n + 1] * 3)
    d = orc_desc(3, [n + 1] * 3, 1, [(PER, PER)] * 3, ext=(1, 1))
    xc = np.arange(n) / n
    Z, Y, X = np.meshgrid(xc, xc, xc, indexing="ij")
    u0 = np.zeros((d.extent[2], d.extent[1], d.extent[0]))
    u0[1:-1, 1:-1, 1:-1] = np.cos(2 * np.pi * (a * X + b * Y + c * Z))
    u0 = oracle.update_padding(d, mesh, u0)
    h = 1.0 / n
    u = oracle.ftcs(d, mesh, u0, 0.2 * h * h, steps)
    lam = 1 - 4 * 0.2 * sum(np.sin(np.pi * k * h) ** 2 for k in (a, b, c))
    exp = lam ** steps * u0[1:-1, 1:-1, 1:-1]
    assert np.max(np.abs(u[1:-1, 1:-1, 1:-1] - exp)) <= 1e-13


def test_ftcs_bruteforce_tiny():
    # 4x4 Neumann/Dirichlet mix against a scalar loop written from the FTCS formula
    n = 4
    mesh = oracle.Mesh([n + 1, n + 1])
    d = orc_desc(2, [n + 1, n + 1], 1, [(NEUM, DIRC), (DIRC, NEUM)], [(0.5, 1.0), (-1.0, 0.25)], ext=(1, 1))
    rng = np.random.default_rng(1)
    u0 = np.zeros((d.extent[1], d.extent[0]))
    u0[1:-1, 1:-1] = rng.uniform(-1, 1, (n, n))
    u0 = oracle.update_padding(d, mesh, u0)
    h = 1.0 / n
    coef = 0.2 * h * h
    u = oracle.ftcs(d, mesh, u0, coef, 1)
    # brute force: ghosts by hand (centre rules: Dirichlet g = 2*bc - u_m; Neumann g = u_m + bc*h signed)
    g = u0.copy()
    exp = np.zeros((n, n))
    for j in range(n):
        for i in range(n):
            c = g[j + 1, i + 1]
            lap = (g[j + 1, i + 2] - 2 * c + g[j + 1, i]) / h ** 2 + (g[j + 2, i + 1] - 2 * c + g[j, i + 1]) / h ** 2
            exp[j, i] = c + coef * lap
    assert np.max(np.abs(u[1:-1, 1:-1] - exp)) <= 1e-12 * np.max(np.abs(exp))
    # ghost rules on the result
    assert u[1:-1, 0] == pytest.approx(u[1:-1, 1] + 0.5 * (-h))      # x start Neumann 0.5
    assert u[1:-1, -1] == pytest.approx(2 * 1.0 - u[1:-1, -2])       # x end Dirichlet 1
    assert u[0, 1:-1] == pytest.approx(2 * -1.0 - u[1, 1:-1])        # y start Dirichlet -1
    assert u[-1, 1:-1] == pytest.approx(u[-2, 1:-1] + 0.25 * h)      # y end Neumann 0.25


def test_padding_pass_replay_matches_closed_form():
    """The product's single-kernel ghost fill uses a closed form of the ordered
    passes; check that closed form (restated here in numpy) against the oracle's
    pass-by-pass replay on random data over many BC/location/ext combinations."""
    rng = np.random.default_rng(5)
    types = [DIRC, NEUM, PER, SYMM, ASYMM]
    for trial in range(60):
        dim = int(rng.integers(1, 4))
        nodes = [int(rng.integers(4, 8)) for _ in range(dim)]
        loc = [int(rng.integers(0, 2)) for _ in range(dim)]
        bc, bcv, ext = [], [], []
        for q in range(dim):
            if rng.random() < 0.3:
                bc.append((PER, PER))
            else:
                bc.append((int(rng.choice(types[:2] + types[3:])), int(rng.choice(types[:2] + types[3:]))))
            bcv.append((float(rng.uniform(-1, 1)), float(rng.uniform(-1, 1))))
            e = int(rng.integers(1, 3))
            ext.append((e, e))
        mesh = oracle.Mesh(nodes)
        d = desc(dim, nodes, loc, bc, bcv, ext=ext)
        d = oracle.finalize(d, mesh)
        u = rng.uniform(-1, 1, tuple(d.extent[q] for q in reversed(range(dim))))
        ref = oracle.update_padding(d, mesh, u.copy())
        got = _closed_form_padding(d, mesh, ref.copy(), u)
        np.testing.assert_array_equal(got, ref)


def _closed_form_padding(d, mesh, interior_src, raw):
    """numpy restatement of the device ghost kernel's closed form."""
    dim = d.dim
    out = raw.copy()
    # step 0 (corner Dirichlet boundary nodes) is part of the interior after the oracle ran
    ls = [d.local_start[q] for q in range(dim)]
    le = [d.local_end[q] for q in range(dim)]
    gs = [d.logical_start[q] for q in range(dim)]
    ge = [d.logical_end[q] for q in range(dim)]
    off = [d.offset[q] for q in range(dim)]

    def at(arr, p):
        return arr[tuple(p[q] - off[q] for q in reversed(range(dim)))]

    it = np.ndindex(*[ge[q] - gs[q] for q in range(dim)])
    for t in it:
        p = [gs[q] + t[q] for q in range(dim)]
        if all(ls[q] <= p[q] < le[q] for q in range(dim)):
            out[tuple(p[q] - off[q] for q in reversed(range(dim)))] = at(interior_src, p)
            continue
        src = list(p)
        for q in range(dim):
            if d.bc_type[q][0] == PER:
                L_ = le[q] - ls[q]
                if src[q] < ls[q]:
                    src[q] += L_
                elif src[q] >= le[q]:
                    src[q] -= L_
        mir = list(src)
        rules = []
        for q in range(dim):
            if ls[q] <= src[q] < le[q]:
                continue
            side = 0 if src[q] < ls[q] else 1
            c = src[q]
            corner = d.loc[q] == 0
            if side == 0:
                mq = 2 * ls[q] - c if corner else 2 * ls[q] - 1 - c
            else:
                mq = 2 * le[q] - 2 - c if corner else 2 * le[q] - 1 - c
            mir[q] = mq
            rules.append((q, side, c, mq))
        v = at(interior_src, mir)
        for (q, side, c, mq) in rules:
            t_ = d.bc_type[q][side]
            g = d.bc_value[q][side]
            corner = d.loc[q] == 0
            if t_ == DIRC:
                if corner:
                    x1 = mesh.x(q, ls[q] if side == 0 else le[q] - 1)
                    x2, x = mesh.x(q, mq), mesh.x(q, c)
                else:
                    x1 = mesh.x(q, ls[q] if side == 0 else le[q])
                    x2 = mesh.x(q, mq) + mesh.dx(q, mq) / 2.
                    x = mesh.x(q, c) + mesh.dx(q, c) / 2.
                A, B = x1 - x, x2 - x
                v = (A * v - B * g) / (A - B)
            elif t_ == NEUM:
                k = (mesh.x(q, c) - mesh.x(q, mq)) if corner else \
                    (mesh.x(q, c) + mesh.dx(q, c) / 2. - mesh.x(q, mq) - mesh.dx(q, mq) / 2.)
                v = v + g * k
            elif t_ == ASYMM:
                v = -v
        out[tuple(p[q] - off[q] for q in reversed(range(dim)))] = v
    return out
