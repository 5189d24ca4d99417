"""The implicit front-end on the GPU: Solve(...) / makeEqnSolveHandler with
device StencilPad extraction, mirroring the reference's DircEqnTest,
NeumEqnTest and PeriodicEqnTest (discrete manufactured solutions) and the
CSRMatrixGenerator golden systems."""
import math

import numpy as np
import pytest

import opflow_amd as of
from opflow_amd import BCType as BC, DimPos as DP, LocOnMesh as LM

pytestmark = pytest.mark.gpu
D1 = of.D1FirstOrderCentered
Cen2Cor = of.d1IntpCenterToCorner


def smooth_heaviside(eps, d):  # Math::smoothHeviside (Math/Function/Numeric.hpp:86-92)
    return np.where(d < -eps, 0.0, np.where(d > eps, 1.0, 0.5 * (1 + d / eps + 1 / math.pi * np.sin(d * math.pi / eps))))


def _field(ctx, mesh, bc, val=0.0, name=""):
    b = of.ExprBuilder().setMesh(mesh).setName(name).setLoc([LM.Center, LM.Center]).setExt(1)
    for d in range(2):
        if bc == BC.Periodic:
            b.setBC(d, DP.start, bc).setBC(d, DP.end, bc)
        else:
            b.setBC(d, DP.start, bc, val).setBC(d, DP.end, bc, val)
    return b.build()


def _setup(ctx, bc, hi, ptrue, rin, rout, p_bc_val=0.0, r_bc_val=1.0):
    mesh = of.MeshBuilder(ctx).newMesh(65 if bc == BC.Dirc else 33, 65 if bc == BC.Dirc else 33) \
        .setMeshOfDim(0, 0., hi).setMeshOfDim(1, 0., hi).build()
    p = _field(ctx, mesh, bc, p_bc_val, "p")
    r = _field(ctx, mesh, bc, r_bc_val, "r")
    p_true = p.clone("ptrue")
    b = p.clone("b")
    p_true.initBy(ptrue)
    dx00 = mesh.dx(0, 0)

    def rfn(x):
        dist = np.sqrt((x[0] - 0.5) * (x[0] - 0.5) + (x[1] - 0.5) * (x[1] - 0.5))
        hevi = smooth_heaviside(dx00 * 8, dist - 0.2)
        return 1. * hevi + (1. - hevi) * rout

    r.initBy(rfn)
    b.assign(of.dx(D1, of.dx(D1, p_true) / Cen2Cor(0, r)) + of.dy(D1, of.dy(D1, p_true) / Cen2Cor(1, r)))
    p.assign(0.0)

    def eqn(e):
        return b == of.dx(D1, of.dx(D1, e) / Cen2Cor(0, r)) + of.dy(D1, of.dy(D1, e) / Cen2Cor(1, r))
    return mesh, p, p_true, eqn


def _check(p, p_true, rel):
    a, t = p.numpy(), p_true.numpy()
    err = np.abs(a[1:-1, 1:-1] - t[1:-1, 1:-1]) / np.abs(t[1:-1, 1:-1])
    return float(np.max(err))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_dirc_eqn_manual_unified_solve(ctx, precision):
    # DircEqnTest.cpp:117-126: tol 1e-10 with PFMG, pointwise relative error <= 1e-10, r jumps 1 -> 1000
    mesh, p, pt, eqn = _setup(ctx, BC.Dirc, 1.0, lambda x: x[0] * (1. - x[0]) * x[1] * (1. - x[1]), None, 1000)
    st = of.Solve(eqn, p, of.StructSolverParams(tol=1e-10, precision=precision),
                  of.StructSolverParams(type=of.StructSolverType.PFMG))
    err = _check(p, pt, 1e-10)
    print(f"Dirc {precision}: iters {st.niter} relerr {st.relerr:.2e} max pointwise rel err {err:.2e}", flush=True)
    assert st.relerr <= 1e-10
    # the 1e-10 tolerance bounds the relative residual; the pointwise error of the
    # solution is larger by up to the operator's condition number, hence 1e-8 here
    assert err <= 1e-8
    # solving twice through a handler (HandlerSolveTwice pattern)
    h = of.makeEqnSolveHandler(eqn, p, of.PrecondStructSolver(of.StructSolverParams(tol=1e-10, staticMat=True)))
    p.assign(0.0)
    st1 = h.solve()
    p.assign(0.0)
    st2 = h.solve()
    assert st1.relerr <= 1e-10 and st2.relerr <= 1e-10
    h.close()


def test_neum_eqn_pinned(ctx):
    # NeumEqnTest.cpp:114-126: pinValue, mean removed, pointwise 5e-8
    mesh, p, pt, eqn = _setup(ctx, BC.Neum, 2 * math.pi, lambda x: np.cos(x[0]) * np.cos(x[1]), None, 1)
    ave = of.rangeReduce(pt, "sum") / (32 * 32)
    pt.assign(pt - ave)
    st = of.Solve(eqn, p, of.StructSolverParams(tol=1e-10, pinValue=True))
    a = p.numpy()
    assert a[1, 1] == 0.0  # the first unknown is pinned at 0 (HYPREEqnSolveHandler.hpp:145-163)
    # the reference's spelling (NeumEqnTest.cpp:121-124): a binary operator and an index functor
    ave = of.rangeReduce(p.assignableRange, lambda a, b: a + b, lambda idx: p[idx]) / (32 * 32)
    assert ave == of.rangeReduce(p, "sum") / (32 * 32)
    p.assign(p - ave)
    err = _check(p, pt, 5e-8)
    print(f"Neum pinned: iters {st.niter} relerr {st.relerr:.2e} err {err:.2e}", flush=True)
    assert err <= 5e-8
    # an expression functor and max / min operators, against numpy on the downloaded values
    diff = p.numpy()[1:-1, 1:-1] - pt.numpy()[1:-1, 1:-1]
    assert of.rangeReduce(p.assignableRange, max, lambda i: p[i] - pt[i]) == diff.max()
    assert of.rangeReduce(p.assignableRange, min, lambda i: p[i] - pt[i]) == diff.min()
    with pytest.raises(of._lib.OpfError):
        of.rangeReduce(p.assignableRange, lambda a, b: a * b, lambda i: p[i])


def test_solver_params_refused_or_honoured(ctx):
    """HYPRE knobs that would change the stopping test or the smoother kind are refused;
    the PFMG settings of the reference's examples (LidDriven2D.cpp:47-48) are accepted."""
    mesh, p, pt, eqn = _setup(ctx, BC.Dirc, 1.0, lambda x: x[0] * (1 - x[0]) * x[1] * (1 - x[1]), None, 1000)
    for bad in (of.StructSolverParams(tol=1e-10, absTol=1e-12), of.StructSolverParams(tol=1e-10, twoNorm=0),
                of.StructSolverParams(tol=1e-10, relChange=1)):
        with pytest.raises(of._lib.OpfError) as ei:
            of.Solve(eqn, p, bad)
        assert ei.value.code == of._lib.OPF_ERR_UNSUPPORTED
    with pytest.raises(of._lib.OpfError):
        of.Solve(eqn, p, of.StructSolverParams(tol=1e-10), of.StructSolverParams(type=3, relaxType=2))
    pf = of.StructSolverParams(type=3, useZeroGuess=True, relaxType=1, rapType=0, numPreRelax=1, numPostRelax=1,
                               skipRelax=0)
    p.assign(5.0)  # useZeroGuess on the solver: starts from 0 whatever p holds
    st = of.Solve(eqn, p, of.StructSolverParams(tol=1e-10, useZeroGuess=True, twoNorm=1), pf)
    assert st.relerr <= 1e-10
    assert _check(p, pt, 1e-8) <= 1e-8


def test_periodic_eqn(ctx):
    # PeriodicEqnTest.cpp:110-121: pointwise 1e-9
    mesh, p, pt, eqn = _setup(ctx, BC.Periodic, 2 * math.pi, lambda x: np.sin(x[0]) * np.cos(x[1]), None, 1)
    st = of.Solve(eqn, p, of.StructSolverParams(tol=1e-10))
    err = _check(p, pt, 1e-9)
    print(f"Periodic: iters {st.niter} relerr {st.relerr:.2e} err {err:.2e}", flush=True)
    assert err <= 1e-9


def test_simple_poisson_golden(ctx):
    # CSRMatrixGeneratorTest SimplePoisson system solved through Solve(): x(0,0) = -0.375, x(1,1) = -1.125
    mesh = of.MeshBuilder(ctx).newMesh(5, 5).setMeshOfDim(0, 0., 4.).setMeshOfDim(1, 0., 4.).build()
    p = _field(ctx, mesh, BC.Dirc, 0.0, "p")
    st = of.Solve(lambda e: 1.0 == of.d2x(of.D2SecondOrderCentered, e) + of.d2y(of.D2SecondOrderCentered, e), p,
                  of.StructSolverParams(tol=1e-13))
    a = p.numpy()
    assert a[1, 1] == pytest.approx(-0.375, abs=1e-12)
    assert a[2, 2] == pytest.approx(-1.125, abs=1e-12)
    assert a[0, 1] == pytest.approx(0.375, abs=1e-12)  # ghost = -mirror after updatePadding


@pytest.mark.parametrize("name,bc,pin", [("SimplePoisson", BC.Dirc, False), ("SimplePoisson_Neum", BC.Neum, True),
                                         ("SimplePoisson_Periodic", BC.Periodic, True)])
def test_device_csr_matches_golden(ctx, name, bc, pin):
    """CSRMatrixGeneratorTest.cpp:57-161: the device-extracted system in CSRMatrixGenerator's layout
    (pin last) equals the reference's golden ptr / col / val / rhs arrays exactly."""
    import json
    import os
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "csr_simple_poisson.json")))["cases"][name]
    mesh = of.MeshBuilder(ctx).newMesh(5, 5).setMeshOfDim(0, 0., 4.).setMeshOfDim(1, 0., 4.).build()
    p = _field(ctx, mesh, bc, 0.0, "p")
    h = of.makeEqnSolveHandler(
        lambda e: of.Equation(1.0, of.d2x(of.D2SecondOrderCentered, e) + of.d2y(of.D2SecondOrderCentered, e)), p,
        of.PrecondStructSolver(of.StructSolverParams(pinValue=pin)))
    ptr, col, val, rhs = h.csr()
    h.close()
    np.testing.assert_array_equal(ptr, gold["ptr"])
    np.testing.assert_array_equal(col, gold["col"])
    np.testing.assert_array_equal(val, gold["val"])
    np.testing.assert_array_equal(rhs, gold["rhs"])


@pytest.mark.parametrize("bc,gval", [(BC.Dirc, 1.5), (BC.Neum, 0.0)], ids=["dirichlet-g1.5", "neumann-pinned"])
def test_static_matrix_rhs_only_solve_matches_full_evaluation(ctx, bc, gval):
    """A staticMat handler re-evaluates only the right-hand side after its first solve
    (generateb, HYPREEqnSolveHandler.hpp:190-209): after the right-hand-side field
    changes, its solve equals a fresh handler's, which evaluates the full StencilPads
    (non-zero boundary values, a variable coefficient in the operator)."""
    mesh = of.MeshBuilder(ctx).newMesh(65, 65).setMeshOfDim(0, 0., 1.).setMeshOfDim(1, 0., 1.).build()
    p = _field(ctx, mesh, bc, gval, "p")
    r = _field(ctx, mesh, bc, 1.0, "r")
    f = _field(ctx, mesh, bc, 0.0, "f")
    r.initBy(lambda x: 1.0 + 9.0 * x[0] * x[1])
    f.initBy(lambda x: np.sin(3 * x[0]) * np.cos(2 * x[1]) + x[0])

    def eqn(e):
        return f == of.dx(D1, of.dx(D1, e) / Cen2Cor(0, r)) + of.dy(D1, of.dy(D1, e) / Cen2Cor(1, r))
    params = of.StructSolverParams(tol=1e-11, staticMat=True, pinValue=bc == BC.Neum)
    h = of.makeEqnSolveHandler(eqn, p, of.PrecondStructSolver(params))
    h.solve()  # first solve: full evaluation, operator built
    f.assign(f * 2.0 + 0.25)  # a new right-hand side: the handler now evaluates only the bias
    p.assign(0.0)
    st_rhs = h.solve()
    got = p.numpy()
    h.close()
    p2 = p.clone("p2")
    p2.assign(0.0)
    h2 = of.makeEqnSolveHandler(eqn, p2, of.PrecondStructSolver(params))
    st_full = h2.solve()
    ref = p2.numpy()
    h2.close()
    print(f"{bc}: rhs-only iters {st_rhs.niter} relerr {st_rhs.relerr:.2e}; full iters {st_full.niter}", flush=True)
    sl = (slice(1, -1), slice(1, -1))
    assert np.max(np.abs(got[sl] - ref[sl])) <= 1e-8 * np.max(np.abs(ref[sl]))


def _csr_relres(h, p):
    """||rhs - A x|| / ||rhs|| of the handler's system (CSRMatrixGenerator rows, assembled
    on the host from the device pads) at the target's current interior values."""
    import scipy.sparse as sp
    ptr, col, val, rhs = h.csr()
    A = sp.csr_matrix((val, col, ptr), shape=(len(rhs), len(rhs)))
    x = p.numpy()[1:-1, 1:-1].ravel()
    return float(np.linalg.norm(rhs - A @ x) / np.linalg.norm(rhs))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nonstatic_rebuild_follows_changed_coefficients(ctx, precision):
    """A handler without staticMat rebuilds its operator on every solve (generateAb,
    HYPREEqnSolveHandler.hpp:125-166), on the device: solves over a variable coefficient,
    the coefficient changed (same layout: the hierarchy's buffers are reused), made
    uniform (constant mode) and variable again; each solve meets the tolerance on the
    system the CSR export assembles on the host from the pads, and equals a fresh
    handler's solve of the same system."""
    mesh = of.MeshBuilder(ctx).newMesh(65, 65).setMeshOfDim(0, 0., 1.).setMeshOfDim(1, 0., 1.).build()
    p = _field(ctx, mesh, BC.Dirc, 0.5, "p")
    r = _field(ctx, mesh, BC.Neum, 0.0, "r")  # mirrored ghosts: a uniform r gives uniform faces
    f = _field(ctx, mesh, BC.Dirc, 0.0, "f")
    f.initBy(lambda x: np.sin(3 * x[0]) * np.cos(2 * x[1]) + x[0])

    def eqn(e):
        return f == of.dx(D1, of.dx(D1, e) / Cen2Cor(0, r)) + of.dy(D1, of.dy(D1, e) / Cen2Cor(1, r))
    params = of.StructSolverParams(tol=1e-10, staticMat=False, precision=precision)
    h = of.makeEqnSolveHandler(eqn, p, of.PrecondStructSolver(params))
    coefs = [lambda x: 1.0 + 9.0 * x[0] * x[1], lambda x: 2.0 + np.cos(5 * x[0]) * x[1],
             lambda x: 0.0 * x[0] + 3.0, lambda x: 1.0 + x[0] * x[0]]
    for k, fn in enumerate(coefs):
        r.initBy(fn)
        r.updatePadding()
        p.assign(0.0)
        st = h.solve()
        got = p.numpy()
        rr = _csr_relres(h, p)
        p2 = p.clone(f"p2_{k}")
        p2.assign(0.0)
        h2 = of.makeEqnSolveHandler(eqn, p2, of.PrecondStructSolver(params))
        st2 = h2.solve()
        ref = p2.numpy()
        h2.close()
        print(f"{precision} coefficient {k}: iters {st.niter} relerr {st.relerr:.2e} csr relres {rr:.2e}; "
              f"fresh iters {st2.niter}", flush=True)
        assert st.relerr <= 1e-10 and rr <= 2e-10
        sl = (slice(1, -1), slice(1, -1))
        assert np.max(np.abs(got[sl] - ref[sl])) <= 1e-8 * np.max(np.abs(ref[sl]))
    h.close()


@pytest.mark.parametrize("bc", [BC.Dirc, BC.Neum, BC.Periodic], ids=["dirc", "neum-pinned", "periodic"])
def test_pads_compiled_equal_interpreter(ctx, bc):
    """The compiled StencilPad kernel (jit.hip) and the pad interpreter give the same system bit
    for bit (the CSR export evaluates the pads on the device): a variable coefficient, a
    function of a coefficient field (sqrt), a reaction term and an advection term."""
    mesh = of.MeshBuilder(ctx).newMesh(41, 29).setMeshOfDim(0, 0., 1.).setMeshOfDim(1, 0., 0.7).build()
    p = _field(ctx, mesh, bc, 0.5, "p")
    r = _field(ctx, mesh, BC.Neum if bc != BC.Periodic else bc, 0.0, "r")
    f = _field(ctx, mesh, bc, 0.0, "f")
    r.initBy(lambda x: 1.0 + 9.0 * x[0] * x[1])
    f.initBy(lambda x: np.sin(3 * x[0]) * np.cos(2 * x[1]))

    def eqn(e):
        return f == (of.dx(D1, of.dx(D1, e) / Cen2Cor(0, r)) + of.dy(D1, of.dy(D1, e) * of.sqrt(Cen2Cor(1, r)))
                     - 2.0 * e + 0.3 * of.dx(of.D1FirstOrderBiasedUpwind, e))
    params = of.StructSolverParams(tol=1e-10, pinValue=bc != BC.Dirc)
    h = of.makeEqnSolveHandler(eqn, p, of.PrecondStructSolver(params))
    out = []
    for mode in (0, 2):
        ctx.set_jit(mode)
        out.append(h.csr())
    ctx.set_jit(1)
    h.close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)
