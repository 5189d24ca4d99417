"""GPU parity of the matrix-free GMG-FCG solver (FP64 and FP32-inner/FP64-IR)
against the oracle (CPU restatement + AMGCL-pinned) and the reference's golden
solution values."""
import numpy as np
import pytest

import opflow_amd as of
import oracle

pytestmark = pytest.mark.gpu
D, N_, P = 1, 2, 3


def _solve(ctx, n, bc, b, precision="fp64", tol=1e-10, sigma=0.0, h=None, maxit=200):
    dim = len(n)
    h = h or [1.0 / n[0]] * dim
    desc = of.poisson_desc(dim, n, h, bc, sigma)
    g = of.PoissonGMG(ctx, desc, precision)
    N = int(np.prod(n))
    bb, xx = of.DeviceBuffer(ctx, N), of.DeviceBuffer(ctx, N)
    bb.upload(b)
    st = g.solve(bb.ptr, xx.ptr, tol=tol, max_iter=maxit)
    x = xx.numpy()
    g.close(); bb.close(); xx.close()
    return x, st


def test_golden_4x4(ctx):
    # CSRMatrixGeneratorTest/AMGCLTest system: x[0] = -0.375, x[5] = -1.125
    x, st = _solve(ctx, [4, 4], [(D, D), (D, D)], -np.ones(16), tol=1e-13, h=[1.0, 1.0])
    assert x[0] == pytest.approx(-0.375, abs=1e-12)
    assert x[5] == pytest.approx(-1.125, abs=1e-12)


def test_apply_matches_oracle(ctx):
    for n, bc in [([40, 24], [(D, N_), (P, P)]), ([16, 8, 12], [(D, D), (N_, D), (P, P)])]:
        dim = len(n)
        h = [1.0 / n[0]] * dim
        g = of.PoissonGMG(ctx, of.poisson_desc(dim, n, h, bc, 0.5), "fp64")
        N = int(np.prod(n))
        x = np.random.default_rng(1).uniform(-1, 1, N)
        xb, yb = of.DeviceBuffer(ctx, N), of.DeviceBuffer(ctx, N)
        xb.upload(x)
        g.apply(xb.ptr, yb.ptr)
        y = yb.numpy()
        ref = oracle.poisson_apply(n, h, bc, x, sigma=0.5)
        np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12 * np.max(np.abs(ref)))
        g.close(); xb.close(); yb.close()


CASES = [([64, 64], [(D, D), (D, D)]), ([64, 32], [(D, N_), (N_, D)]), ([32, 32], [(P, P), (D, D)]),
         ([32, 32, 32], [(D, D)] * 3), ([16, 32, 16], [(N_, D), (P, P), (D, N_)])]


@pytest.mark.parametrize("n,bc", CASES, ids=[f"{len(c[0])}d-{i}" for i, c in enumerate(CASES)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_matches_oracle(ctx, n, bc, precision):
    N = int(np.prod(n))
    b = np.random.default_rng(7).normal(size=N)
    x, st = _solve(ctx, n, bc, b, precision, tol=1e-10)
    print(f"{n} {bc} {precision}: iters {st.niter} refine {st.nrefine} relres {st.relerr:.2e}", flush=True)
    assert st.converged and st.relerr <= 1e-10
    ref, it, rel = oracle.pcg_mg(n, [1.0 / n[0]] * len(n), bc, b, tol=1e-12)
    assert rel <= 1e-12
    err = np.max(np.abs(x - ref)) / np.max(np.abs(ref))
    assert err <= 1e-7, err


def test_solve_singular_neumann_and_periodic(ctx):
    for n, bc in [([64, 64], [(N_, N_), (N_, N_)]), ([32, 32], [(P, P), (P, P)]), ([16, 16, 16], [(N_, N_), (P, P), (N_, N_)])]:
        N = int(np.prod(n))
        b = np.random.default_rng(8).normal(size=N)
        b -= b.mean()  # compatible right-hand side
        for precision in ("fp64", "mixed"):
            x, st = _solve(ctx, n, bc, b, precision, tol=1e-10)
            assert st.relerr <= 1e-10, (n, bc, precision, st)
            # the residual check is on the original system; compare to the oracle up to a constant
            r = b - oracle.poisson_apply(n, [1.0 / n[0]] * len(n), bc, x)
            assert np.linalg.norm(r) <= 1.01e-10 * np.linalg.norm(b)


def test_heat_backward_euler_sigma(ctx):
    n = [128, 128]
    h = 1.0 / 128
    sigma = 1.0 / (0.5 * h * h)  # 1/(alpha dt)
    b = np.random.default_rng(9).normal(size=128 * 128)
    x, st = _solve(ctx, n, [(D, D), (D, D)], b, "mixed", tol=1e-10, sigma=sigma)
    r = b - oracle.poisson_apply(n, [h, h], [(D, D), (D, D)], x, sigma=sigma)
    assert np.linalg.norm(r) <= 1.01e-10 * np.linalg.norm(b)
    assert st.niter < 30


# 3-D grids of >= 2^18 cells with n0 >= 64 run the z-pipelined LDS-ring kernels
# (boundary conditions folded into the halo). 80 x 40 and 144 x 24 leave
# partial 32 x 16 (FP64) and 64 x 8 (FP32) tiles; every extent coarsens to a
# dense coarsest level of at most 1024 cells.
ZP_APPLY = [([80, 40, 88], [(D, N_), (P, P), (N_, D)]), ([144, 24, 80], [(P, P), (D, D), (D, N_)]),
            ([96, 48, 64], [(N_, N_), (D, D), (P, P)])]


@pytest.mark.parametrize("n,bc", ZP_APPLY, ids=["ragged-xy", "ragged-x", "full-tiles"])
def test_zpipe_apply_matches_oracle(ctx, n, bc):
    h = [1.0 / n[0]] * 3
    g = of.PoissonGMG(ctx, of.poisson_desc(3, n, h, bc, 0.5), "fp64")
    N = int(np.prod(n))
    x = np.random.default_rng(11).uniform(-1, 1, N)
    xb, yb = of.DeviceBuffer(ctx, N), of.DeviceBuffer(ctx, N)
    xb.upload(x)
    g.apply(xb.ptr, yb.ptr)
    y = yb.numpy()
    ref = oracle.poisson_apply(n, h, bc, x, sigma=0.5)
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12 * np.max(np.abs(ref)))
    g.close(); xb.close(); yb.close()


ZP_SOLVE = [([64, 64, 64], [(D, D)] * 3), ([96, 48, 64], [(D, N_), (P, P), (N_, D)]),
            ([80, 40, 88], [(D, D), (D, N_), (P, P)]), ([144, 24, 80], [(D, D), (N_, D), (D, D)])]


@pytest.mark.parametrize("n,bc", ZP_SOLVE, ids=["64^3-dirichlet", "mixed-bc", "ragged-xy", "ragged-x"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_zpipe_solve_matches_oracle(ctx, n, bc, precision):
    N = int(np.prod(n))
    b = np.random.default_rng(12).normal(size=N)
    x, st = _solve(ctx, n, bc, b, precision, tol=1e-10)
    print(f"{n} {bc} {precision}: iters {st.niter} refine {st.nrefine} relres {st.relerr:.2e}", flush=True)
    assert st.converged and st.relerr <= 1e-10
    assert st.niter <= 20  # a V(2,2) Jacobi cycle with the halo-folded boundary diagonal keeps its rate
    ref, it, rel = oracle.pcg_mg(n, [1.0 / n[0]] * 3, bc, b, tol=1e-12)
    assert rel <= 1e-12
    err = np.max(np.abs(x - ref)) / np.max(np.abs(ref))
    assert err <= 1e-7, err


def test_zpipe_singular(ctx):
    n, bc = [64, 32, 128], [(P, P), (N_, N_), (P, P)]
    N = int(np.prod(n))
    b = np.random.default_rng(13).normal(size=N)
    b -= b.mean()
    for precision in ("fp64", "mixed"):
        x, st = _solve(ctx, n, bc, b, precision, tol=1e-10)
        assert st.relerr <= 1e-10, (precision, st)
        r = b - oracle.poisson_apply(n, [1.0 / n[0]] * 3, bc, x)
        assert np.linalg.norm(r) <= 1.01e-10 * np.linalg.norm(b)


@pytest.mark.parametrize("n", [[2048, 2048], [256, 256, 256]], ids=["2048^2", "256^3"])
def test_full_size_fp64_vs_mixed(ctx, n):
    """Implicit config at full size: bumps + 1e-3 noise RHS, Dirichlet; both precisions reach 1e-10,
    with the residual ||b - A x|| / ||b|| recomputed by the oracle's operator, agree with each other,
    and the FP64 solution agrees with the oracle's own solve (flexible PCG + V-cycle) to 1e-7."""
    dim = len(n)
    mb = of.MeshBuilder(ctx).newMesh(*[k + 1 for k in n])
    mesh = mb.build()
    b_ = of.ExprBuilder().setMesh(mesh).setLoc(of.LocOnMesh.Center).setExt(1)
    for d in range(dim):
        b_.setBC(d, 0, of.BCType.Dirc, 0.0).setBC(d, 1, of.BCType.Dirc, 0.0)
    f = b_.build()
    rhs = of.gen_bumps(f, 0xC0FFEE + 3, 4, 1e-3)
    sl = tuple(slice(1, -1) for _ in range(dim))
    b = np.ascontiguousarray(rhs[sl]).ravel()
    f.close()
    x64, s64 = _solve(ctx, n, [(D, D)] * dim, b, "fp64", tol=1e-10)
    x32, s32 = _solve(ctx, n, [(D, D)] * dim, b, "mixed", tol=1e-10)
    print(f"{n}: fp64 iters {s64.niter} relres {s64.relerr:.2e}; mixed iters {s32.niter} refine {s32.nrefine} "
          f"relres {s32.relerr:.2e}", flush=True)
    assert s64.relerr <= 1e-10 and s32.relerr <= 1e-10
    assert np.max(np.abs(x64 - x32)) <= 1e-8 * np.max(np.abs(x64))
    h = [1.0 / n[0]] * dim
    bcs = [(D, D)] * dim
    nb = np.linalg.norm(b)
    for name, x in (("fp64", x64), ("mixed", x32)):
        rel = np.linalg.norm(b - oracle.poisson_apply(n, h, bcs, x)) / nb
        print(f"{n} {name}: oracle-recomputed relres {rel:.3e}", flush=True)
        assert rel <= 1.01e-10
    # (the oracle's FP64 residual stagnates near 1e-11..1e-12 at these sizes and its stopping test
    # uses the recurrence residual, so it is asked for 1e-11 and checked at the configuration's
    # 1e-10; the two solutions agree to ~6e-12)
    ref, it, rel_ref = oracle.pcg_mg(n, h, bcs, b, tol=1e-11)
    err = np.max(np.abs(x64 - ref)) / np.max(np.abs(ref))
    print(f"{n}: oracle solve {it} iterations relres {rel_ref:.2e}; max |x64 - x_oracle| / max|x| {err:.3e}",
          flush=True)
    assert rel_ref <= 1e-10 and err <= 1e-7
