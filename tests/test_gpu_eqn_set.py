"""Equation sets on the GPU: SolveEqns and the multi-equation CSRMatrixGenerator
(UnifiedSolve.hpp:69-100, CSRMatrixGenerator.hpp:26-41), mirroring the reference's
EqnSetTest.cpp (2 and 10 decoupled equations, Dirichlet and pinned Neumann) and
CSRMatrixGeneratorTest.cpp:163-250 (block structure), plus coupled sets checked
against the oracle's assembly (oracle.expr.linear_system) and dense solves."""
import json
import os

import numpy as np
import pytest

import opflow_amd as of
import oracle
from oracle import expr as X
from opflow_amd import BCType as BC, DimPos as DP, LocOnMesh as LM

pytestmark = pytest.mark.gpu
D2 = of.D2SecondOrderCentered
UP = of.D1FirstOrderBiasedUpwind
DOWN = of.D1FirstOrderBiasedDownwind
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _mesh(ctx, nx, ny, hx, hy):
    return of.MeshBuilder(ctx).newMesh(nx + 1, ny + 1).setMeshOfDim(0, 0., hx).setMeshOfDim(1, 0., hy).build()


def _field(mesh, bcs, name):
    """bcs: per dim (start, end) as (type, value)."""
    b = of.ExprBuilder().setMesh(mesh).setName(name).setLoc([LM.Center, LM.Center]).setExt(1)
    for d in range(2):
        for side, (t, v) in zip((DP.start, DP.end), bcs[d]):
            if t == BC.Periodic:
                b.setBC(d, side, t)
            else:
                b.setBC(d, side, t, v)
    return b.build()


def _uni(t, v=0.0):
    return [((t, v), (t, v)), ((t, v), (t, v))]


def _oracle(getters, targets, om, pins=None):
    unk = [of.StencilUnknown(i, t) for i, t in enumerate(targets)]
    eqs = []
    for g in getters:
        eqn = g(*unk)
        nodes, fields, root = of.Expr(3, eqn.lhs, eqn.rhs).flatten()
        eqs.append((nodes[:-1], nodes[root][1], nodes[root][2], [(f.numpy(), f.desc) for f in fields]))
    unknowns = [(t.numpy(), t.desc) for t in targets]
    return X.linear_system(eqs, unknowns, om, lambda i, a: oracle.update_padding(targets[i].desc, om, a), pins)


def _dense(m):
    n = len(m.row) - 1
    A = np.zeros((n, n))
    for r in range(n):
        A[r, m.col[m.row[r]:m.row[r + 1]]] = m.val[m.row[r]:m.row[r + 1]]
    return A


def _interior(f):
    return f.numpy()[1:-1, 1:-1].reshape(-1)


def _lap(e):
    return of.d2x(D2, e) + of.d2y(D2, e)


@pytest.mark.parametrize("name,bct,pin", [("SimplePoisson", BC.Dirc, False), ("SimplePoisson_Neum", BC.Neum, True)])
def test_csr_two_equations_block_structure(ctx, name, bct, pin):
    # CSRMatrixGeneratorTest SimplePoisson_2Eqn / _Neum_2Eqn: targets (p, p); the second
    # block repeats the first 16 columns later; the first equals the golden system
    gold = json.load(open(os.path.join(GOLD, "csr_simple_poisson.json")))["cases"][name]
    p = _field(_mesh(ctx, 4, 4, 4., 4.), _uni(bct), "p")
    st = of.makeStencilHolder(of.makeEqnHolder(
        [lambda e, _: of.Equation(1.0, _lap(e)), lambda _, e: of.Equation(1.0, _lap(e))], [p, p]))
    mapper = of.ColoredMDRangeMapper(p.assignableRange, p.assignableRange)
    m = of.CSRMatrixGenerator.generate(st, mapper, [pin, pin])
    h = len(m.row) // 2
    np.testing.assert_array_equal(m.row[:h + 1], gold["ptr"])
    np.testing.assert_array_equal(m.col[:len(m.col) // 2], gold["col"])
    np.testing.assert_array_equal(m.val[:len(m.val) // 2], gold["val"])
    np.testing.assert_array_equal(m.rhs[:h], gold["rhs"])
    np.testing.assert_array_equal(m.row[h:] - m.row[h], m.row[:h + 1])
    np.testing.assert_array_equal(m.col[len(m.col) // 2:] - 16, m.col[:len(m.col) // 2])
    np.testing.assert_array_equal(m.val[len(m.val) // 2:], m.val[:len(m.val) // 2])
    np.testing.assert_array_equal(m.rhs[h:], m.rhs[:h])


def _coupled_case(ctx, nx=6, ny=5):
    mesh = _mesh(ctx, nx, ny, 1.0, 0.8)
    om = oracle.Mesh([nx + 1, ny + 1], [(0., 1.0), (0., 0.8)])
    u = _field(mesh, [((BC.Dirc, 0.7), (BC.Neum, -0.2)), ((BC.Dirc, 0.0), (BC.Dirc, 1.5))], "u")
    v = _field(mesh, [((BC.Neum, 0.3), (BC.Dirc, -0.4)), ((BC.Neum, 0.0), (BC.Dirc, 0.25))], "v")
    r = _field(mesh, _uni(BC.Dirc, 1.0), "r")
    f = _field(mesh, _uni(BC.Dirc, 0.0), "f")
    r.initBy(lambda x: 1.0 + 2.0 * x[0] * x[1])
    f.initBy(lambda x: np.sin(3 * x[0]) + x[1])
    Cen2Cor = of.d1IntpCenterToCorner
    D1 = of.D1FirstOrderCentered

    def eq0(a, b):  # variable-coefficient diffusion of a, reaction with b
        return of.Equation(f, of.dx(D1, of.dx(D1, a) / Cen2Cor(0, r)) + of.dy(D1, of.dy(D1, a) / Cen2Cor(1, r))
                           - 4.0 * a + 0.5 * r * b)

    def eq1(a, b):  # diffusion of b, advection of a in x and y
        return of.Equation(2.0 * f - 1.0, _lap(b) - 3.0 * b + of.dx(UP, a) + 0.5 * of.dy(DOWN, a))
    return mesh, om, u, v, [eq0, eq1]


def test_csr_coupled_matches_oracle(ctx):
    mesh, om, u, v, eqs = _coupled_case(ctx)
    m = of.CSRMatrixGenerator.generate(of.makeStencilHolder(of.makeEqnHolder(eqs, [u, v])))
    A, rhs = _oracle(eqs, [u, v], om)
    Ad = _dense(m)
    scale = np.max(np.abs(A))
    assert np.max(np.abs(Ad - A)) <= 1e-12 * scale
    np.testing.assert_allclose(m.rhs, rhs, rtol=0, atol=1e-12 * np.max(np.abs(rhs)))
    # no entries beyond what the pads touch: the coupling blocks are diagonal / upwind
    nnz = np.count_nonzero(Ad)
    assert len(m.col) - nnz <= 0.1 * len(m.col)


def test_solve_eqns_coupled_matches_dense_solve(ctx):
    mesh, om, u, v, eqs = _coupled_case(ctx, 8, 8)
    A, rhs = _oracle(eqs, [u, v], om)
    ref = np.linalg.solve(A, rhs)
    params = [of.IJSolverParams(tol=1e-12), of.IJSolverParams()]
    st = of.SolveEqns(eqs[0], eqs[1], u, v, of.ColoredMDRangeMapper(u.assignableRange, v.assignableRange), params)
    got = np.concatenate([_interior(u), _interior(v)])
    print(f"coupled 8x8: iters {st.niter} relerr {st.relerr:.2e}", flush=True)
    assert st.converged and st.relerr <= 1e-12
    assert np.max(np.abs(got - ref)) <= 1e-9 * np.max(np.abs(ref))
    # padding was updated with each target's own boundary rule
    a = u.numpy()
    np.testing.assert_allclose(a[1:-1, 0], 2 * 0.7 - a[1:-1, 1], rtol=0, atol=1e-12)


@pytest.mark.parametrize("bct,pin", [(BC.Dirc, False), (BC.Neum, True)], ids=["dirichlet", "neumann-pinned"])
def test_solve_eqns_two_equations(ctx, bct, pin):
    # EqnSetTest SimplePoisson_2Eqn / SimplePoisson_Neum_2Eqn: p == p2, both equal the single solve
    mesh = _mesh(ctx, 4, 4, 4., 4.)
    om = oracle.Mesh([5, 5], [(0., 4.), (0., 4.)])
    p = _field(mesh, _uni(bct), "p")
    p2 = p.clone("p2")
    p.assign(0.0)
    p2.assign(0.0)
    eqs = [lambda e, _: of.Equation(1.0, _lap(e)), lambda _, e: of.Equation(1.0, _lap(e))]
    params = [of.IJSolverParams(pinValue=pin, tol=1e-13), of.IJSolverParams(pinValue=pin)]
    st = of.SolveEqns(*eqs, p, p2, of.ColoredMDRangeMapper(p.assignableRange, p.assignableRange), params)
    assert st.relerr <= 1e-13
    np.testing.assert_array_equal(p.numpy(), p2.numpy())
    A, rhs = _oracle([lambda e: of.Equation(1.0, _lap(e))], [p], om, [pin])
    ref = np.linalg.solve(A, rhs)
    got = _interior(p)
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))
    if pin:
        assert got[-1] == 0.0  # the pinned (last) unknown
    else:
        assert got[0] == pytest.approx(-0.375, abs=1e-12)  # AMGCLTest x[0]


def test_solve_eqns_ten_equations_batched(ctx):
    # EqnSetTest SimplePoisson_10Eqn: ten copies of one operator are the slots of one
    # hierarchy; every target gets the same solution
    mesh = _mesh(ctx, 32, 32, 1., 1.)
    om = oracle.Mesh([33, 33], [(0., 1.), (0., 1.)])
    ps = [_field(mesh, _uni(BC.Dirc), f"p{i}") for i in range(10)]
    for q in ps:
        q.assign(0.0)

    def getter(i):
        return lambda *e: of.Equation(1.0, _lap(e[i]))
    eqs = [getter(i) for i in range(10)]
    st = of.SolveEqns(*eqs, *ps, of.ColoredMDRangeMapper(*[q.assignableRange for q in ps]),
                      [of.IJSolverParams(tol=1e-12)] + [of.IJSolverParams()] * 9)
    assert st.relerr <= 1e-12
    for q in ps[1:]:
        np.testing.assert_array_equal(q.numpy(), ps[0].numpy())
    ref = oracle.pcg_mg([32, 32], [1 / 32] * 2, [(1, 1), (1, 1)], -np.ones(32 * 32), tol=1e-13)[0]
    got = _interior(ps[0])
    assert np.max(np.abs(got - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_solve_eqns_coupled_at_size_residual(ctx):
    # 128x128 coupled set: the device solution's residual in the CSR system (whose
    # assembly the small cases pin against the oracle) meets the tolerance
    import scipy.sparse as sp
    mesh, om, u, v, eqs = _coupled_case(ctx, 128, 128)
    m = of.CSRMatrixGenerator.generate(of.makeEqnHolder(eqs, [u, v]))
    st = of.SolveEqns(*eqs, u, v, None, [of.IJSolverParams(tol=1e-10), of.IJSolverParams()])
    A = sp.csr_matrix((m.val, m.col, m.row))
    x = np.concatenate([_interior(u), _interior(v)])
    rel = np.linalg.norm(m.rhs - A @ x) / np.linalg.norm(m.rhs)
    print(f"coupled 128^2: iters {st.niter} relerr {st.relerr:.2e} host residual {rel:.2e}", flush=True)
    assert st.relerr <= 1e-10 and rel <= 1.5e-10


def test_solve_eqns_mixed_precision_batch(ctx):
    mesh = _mesh(ctx, 64, 64, 1., 1.)
    ps = [_field(mesh, _uni(BC.Dirc, 0.5 * i), f"p{i}") for i in range(3)]
    eqs = [lambda a, b, c: of.Equation(1.0, _lap(a)), lambda a, b, c: of.Equation(2.0, _lap(b)),
           lambda a, b, c: of.Equation(-1.0, _lap(c))]
    out = []
    for prec in ("fp64", "mixed"):
        for q in ps:
            q.assign(0.0)
        st = of.SolveEqns(*eqs, *ps, None, [of.IJSolverParams(tol=1e-10, precision=prec)] + [of.IJSolverParams()] * 2)
        assert st.relerr <= 1e-10
        out.append(np.concatenate([_interior(q) for q in ps]))
    assert np.max(np.abs(out[0] - out[1])) <= 1e-7 * np.max(np.abs(out[0]))


def test_solve_eqns_rejects_nonlinear(ctx):
    mesh = _mesh(ctx, 8, 8, 1., 1.)
    u = _field(mesh, _uni(BC.Dirc), "u")
    v = _field(mesh, _uni(BC.Dirc), "v")
    with pytest.raises(of._lib.OpfError, match="nonlinear"):
        of.SolveEqns(lambda a, b: of.Equation(1.0, _lap(a) + a * b), lambda a, b: of.Equation(1.0, _lap(b)), u, v,
                     None, None)


def test_solve_eqns_coupled_with_pinned_neumann_block(ctx):
    # a singular (pure Neumann) diagonal block made regular by pinValue, inside a coupled set
    mesh = _mesh(ctx, 8, 8, 1., 1.)
    om = oracle.Mesh([9, 9], [(0., 1.), (0., 1.)])
    a = _field(mesh, _uni(BC.Neum, 0.0), "a")
    b = _field(mesh, _uni(BC.Dirc, 0.5), "b")
    f = _field(mesh, _uni(BC.Dirc, 0.0), "f")
    f.initBy(lambda x: np.cos(3 * x[0]) * x[1])
    eqs = [lambda p, q: of.Equation(f, _lap(p) + 0.3 * q), lambda p, q: of.Equation(1.0, _lap(q) - 2.0 * q + 0.2 * p)]
    A, rhs = _oracle(eqs, [a, b], om, [True, False])
    ref = np.linalg.solve(A, rhs)
    st = of.SolveEqns(*eqs, a, b, None, [of.IJSolverParams(tol=1e-11, pinValue=True), of.IJSolverParams()])
    got = np.concatenate([_interior(a), _interior(b)])
    print(f"coupled pinned: iters {st.niter} relerr {st.relerr:.2e}", flush=True)
    assert st.relerr <= 1e-11
    assert np.max(np.abs(got - ref)) <= 1e-8 * np.max(np.abs(ref))


def _advdiff(f, c=(8.0, 3.0)):
    # steady advection-diffusion: non-symmetric StencilPads (upwind / downwind differences)
    return lambda e: of.Equation(f, _lap(e) - c[0] * of.dx(UP, e) + c[1] * of.dy(DOWN, e))


@pytest.mark.parametrize("bct,pin", [(BC.Dirc, False), (BC.Neum, True)], ids=["dirichlet", "neumann-pinned"])
def test_solve_nonsymmetric_single_equation(ctx, bct, pin):
    """An operator with advection terms (not symmetric) through Solve(): BiCGStab on the
    system HYPREEqnSolveHandler builds (pinValue pins the FIRST unknown), checked against
    a dense solve of the oracle's assembly of the same system."""
    mesh = _mesh(ctx, 24, 20, 1.0, 1.0)
    om = oracle.Mesh([25, 21], [(0., 1.), (0., 1.)])
    p = _field(mesh, [((bct, 0.4), (bct, -0.2)), ((bct, 0.0), (bct, 0.3))] if bct == BC.Dirc else _uni(bct), "p")
    f = _field(mesh, _uni(BC.Dirc, 0.0), "f")
    f.initBy(lambda x: np.sin(5 * x[0]) * x[1] - 0.3 if pin else np.sin(5 * x[0]) * x[1])
    A, rhs = _oracle([_advdiff(f)], [p], om)
    if pin:
        A[0, :] = 0.0
        A[0, 0] = 1.0
        rhs[0] = 0.0
    ref = np.linalg.solve(A, rhs)
    p.assign(0.0)
    st = of.Solve(_advdiff(f), p, of.StructSolverParams(tol=1e-12, pinValue=pin))
    got = _interior(p)
    print(f"non-symmetric {bct}: iters {st.niter} relerr {st.relerr:.2e}", flush=True)
    assert st.relerr <= 1e-12
    assert np.max(np.abs(got - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_solve_eqns_nonsymmetric_block(ctx):
    # a block-diagonal set whose blocks are not symmetric goes to BiCGStab as well
    mesh = _mesh(ctx, 16, 16, 1.0, 1.0)
    om = oracle.Mesh([17, 17], [(0., 1.), (0., 1.)])
    u = _field(mesh, _uni(BC.Dirc, 0.5), "u")
    v = _field(mesh, _uni(BC.Dirc, -0.5), "v")
    f = _field(mesh, _uni(BC.Dirc, 0.0), "f")
    f.initBy(lambda x: x[0] * x[1])
    eqs = [lambda a, b: _advdiff(f)(a), lambda a, b: _advdiff(f, (-4.0, 6.0))(b)]
    A, rhs = _oracle(eqs, [u, v], om)
    ref = np.linalg.solve(A, rhs)
    st = of.SolveEqns(*eqs, u, v, None, [of.IJSolverParams(tol=1e-12), of.IJSolverParams()])
    got = np.concatenate([_interior(u), _interior(v)])
    assert st.relerr <= 1e-12
    assert np.max(np.abs(got - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_solve_nonsymmetric_at_size_residual(ctx):
    import scipy.sparse as sp
    mesh = _mesh(ctx, 256, 256, 1.0, 1.0)
    p = _field(mesh, _uni(BC.Dirc, 0.0), "p")
    f = _field(mesh, _uni(BC.Dirc, 0.0), "f")
    f.initBy(lambda x: np.exp(-20 * ((x[0] - 0.4) ** 2 + (x[1] - 0.6) ** 2)))
    eq = _advdiff(f, (40.0, 25.0))
    h = of.makeEqnSolveHandler(eq, p, of.PrecondStructSolver(of.StructSolverParams(tol=1e-10)))
    ptr, col, val, rhs = h.csr()
    p.assign(0.0)
    st = h.solve()
    h.close()
    A = sp.csr_matrix((val, col, ptr))
    x = _interior(p)
    rel = np.linalg.norm(rhs - A @ x) / np.linalg.norm(rhs)
    print(f"non-symmetric 256^2: iters {st.niter} relerr {st.relerr:.2e} host residual {rel:.2e}", flush=True)
    assert st.relerr <= 1e-10 and rel <= 1.5e-10


def test_solve_eqns_argument_errors(ctx):
    """Sets the engine cannot take are refused with a message, not solved wrongly."""
    mesh = _mesh(ctx, 8, 8, 1.0, 1.0)
    u = _field(mesh, _uni(BC.Dirc), "u")
    v = _field(mesh, _uni(BC.Dirc), "v")
    other = _field(_mesh(ctx, 8, 6, 1.0, 1.0), _uni(BC.Dirc), "w")
    with pytest.raises(of._lib.OpfError, match="mesh|layout"):
        of.SolveEqns(lambda a, b: of.Equation(1.0, _lap(a)), lambda a, b: of.Equation(1.0, _lap(b)), u, other,
                     None, None)
    with pytest.raises(of._lib.OpfError, match="own target"):
        of.SolveEqns(lambda a, b: of.Equation(1.0, _lap(b)), lambda a, b: of.Equation(1.0, _lap(a)), u, v, None, None)
    with pytest.raises(of._lib.OpfError, match="no unknown"):
        of.SolveEqns(lambda a, b: of.Equation(1.0, 2.0), lambda a, b: of.Equation(1.0, _lap(b)), u, v, None, None)
    with pytest.raises(of._lib.OpfError, match="5/7-point"):  # a diagonal neighbour of b
        of.SolveEqns(lambda a, b: of.Equation(1.0, _lap(a) + of.dx(UP, of.dy(UP, b))),
                     lambda a, b: of.Equation(1.0, _lap(b)), u, v, None, None)


def test_pads_compiled_equal_interpreter_set(ctx):
    """The compiled StencilPad kernel (jit.hip) and the pad interpreter build the same coupled
    system bit for bit: every coefficient block and the right-hand side of the set's CSR."""
    mesh, om, u, v, eqs = _coupled_case(ctx, 19, 14)
    out = []
    for mode in (0, 2):
        ctx.set_jit(mode)
        m = of.CSRMatrixGenerator.generate(of.makeStencilHolder(of.makeEqnHolder(eqs, [u, v])))
        out.append((np.array(m.row), np.array(m.col), np.array(m.val), np.array(m.rhs)))
    ctx.set_jit(1)
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)
