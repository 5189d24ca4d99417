"""Continuous-batching server: every served result meets the tolerance, a
request's result is the same alone and in a mix (bitwise: slots never
interact and reductions are fixed-order), graph replays equal direct
launches, and memory-based admission queues requests that do not fit."""
import numpy as np
import pytest

import opflow_amd as of
import oracle

pytestmark = pytest.mark.gpu
D = 1


def _desc(n):
    dim = len(n)
    return of.poisson_desc(dim, n, [1.0 / n[0]] * dim, [(D, D)] * dim)


def _serve(ctx, reqs, budget=8 << 30, batched=True, precision="mixed", late=(), max_inflight=16):
    srv = of.Server(ctx, budget, max_inflight=max_inflight, precision=precision, batched=batched, tol=1e-8)
    outs, ids = {}, {}
    for k, (n, seed, nb) in enumerate(reqs):
        outs[k] = np.zeros(int(np.prod(n)))
        ids[srv.submit(_desc(n), seed, nb, out=outs[k])] = k
    results = {}
    rounds = 0
    late = list(late)
    while len(results) < len(reqs) + len(late) and rounds < 10000:
        srv.step()
        rounds += 1
        if rounds == 3:  # requests arriving while others are on the GPU
            for n, seed, nb in late:
                k = len(outs)
                outs[k] = np.zeros(int(np.prod(n)))
                ids[srv.submit(_desc(n), seed, nb, out=outs[k])] = k
            reqs = list(reqs) + late
            late = []
        for r in srv.poll():
            results[ids[r.id]] = r
        q, f = srv.pending()
        assert f <= max_inflight
    srv.close()
    return outs, results


MIX = [([64, 64], 1, 3), ([128, 128], 2, 5), ([64, 64], 3, 1), ([32, 32, 32], 4, 8), ([256, 256], 5, 2),
       ([64, 64, 64], 6, 4), ([128, 128], 7, 6), ([32, 32, 32], 8, 2)]


def test_served_results_meet_tolerance(ctx):
    outs, res = _serve(ctx, MIX)
    assert len(res) == len(MIX)
    for k, (n, seed, nb) in enumerate(MIX):
        r = res[k]
        assert r.status == 0 and r.state.converged, (n, r.state.relerr)
        b = of.gen_bumps_grid(_desc(n), seed, nb)
        res_vec = b - oracle.poisson_apply(n, [1.0 / n[0]] * len(n), [(D, D)] * len(n), outs[k])
        assert np.linalg.norm(res_vec) <= 1.01e-8 * np.linalg.norm(b)
        assert r.t_done >= r.t_admit >= r.t_submit - 1e-9


@pytest.mark.parametrize("target", [([128, 128], 2, 5), ([64, 64, 64], 9, 4)], ids=["128^2", "64^3"])
def test_alone_equals_in_mix(ctx, target):
    alone, _ = _serve(ctx, [target])
    mixed, _ = _serve(ctx, MIX[:3], late=[target] + MIX[3:])
    k = 3  # first late request
    np.testing.assert_array_equal(alone[0], mixed[k])
    # and within tolerance of a direct (non-served) mixed-precision solve
    desc = _desc(target[0])
    b = of.gen_bumps_grid(desc, target[1], target[2])
    g = of.PoissonGMG(ctx, desc, "mixed")
    bb, xx = of.DeviceBuffer(ctx, b.size), of.DeviceBuffer(ctx, b.size)
    bb.upload(b)
    st = g.solve(bb.ptr, xx.ptr, tol=1e-8)
    x = xx.numpy()
    assert np.max(np.abs(x - alone[0])) <= 1e-6 * np.max(np.abs(x))
    g.close(); bb.close(); xx.close()


def test_graph_replays_match_direct_launches(ctx):
    a, _ = _serve(ctx, MIX, batched=True)
    b, _ = _serve(ctx, MIX, batched=False)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])


def test_memory_admission_queues(ctx):
    """Budget just above one 256^2 group (8 slots at N <= 2^18): the 128^2
    requests cannot be co-resident with it, wait for it to finish and be
    evicted, and still finish to tolerance."""
    desc = _desc([256, 256])
    N = 256 * 256
    group = ctx.lib.opf_gmg_footprint(desc, 1) * 8 + 8 * N * 16 + 4096
    budget = int(group * 1.1)
    reqs = [([256, 256], s, 2) for s in range(3)] + [([128, 128], 10 + s, 3) for s in range(3)]
    outs, res = _serve(ctx, reqs, budget=budget)
    assert len(res) == len(reqs)
    assert all(r.state.converged for r in res.values())
    first_done = min(res[k].t_done for k in range(3))
    assert all(res[k].t_admit >= first_done for k in range(3, 6))
    for k, (n, seed, nb) in enumerate(reqs):
        b = of.gen_bumps_grid(_desc(n), seed, nb)
        r = b - oracle.poisson_apply(n, [1.0 / n[0]] * 2, [(D, D)] * 2, outs[k])
        assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(b)


def test_device_rhs_generator_matches_host(ctx):
    """Device bumps + noise equal the host generator's (same SplitMix64 draws) to the last bits of exp/log/cos."""
    for n, nb, noise in (([96, 80], 5, 0.0), ([24, 20, 16], 8, 1e-3), ([64, 64], 32, 0.5)):
        desc = _desc(n)
        host = of.gen_bumps_grid(desc, 1234 + nb, nb, noise)
        buf = of.DeviceBuffer(ctx, host.size)
        of.gen_bumps_grid_dev(ctx, desc, 1234 + nb, nb, noise, buf)
        dev = buf.numpy()
        buf.close()
        assert np.max(np.abs(dev - host)) <= 1e-14 * np.max(np.abs(host)), (n, nb, noise)


LARGE = [([1024, 1024], 21, 3), ([128, 128, 128], 22, 5), ([2048, 2048], 23, 8), ([256, 256, 256], 24, 2),
         ([64, 64], 25, 1)]


def test_served_large_shapes(ctx):
    """The configuration's largest grids (1024^2, 2048^2, 128^3, 256^3) served in one mix, some
    arriving while others run: every result meets 1e-8 with the residual recomputed by the
    oracle's operator, and the 256^3 request's result equals the one it gets served alone."""
    outs, res = _serve(ctx, LARGE[:2], late=LARGE[2:])
    assert len(res) == len(LARGE)
    for k, (n, seed, nb) in enumerate(LARGE):
        r = res[k]
        assert r.status == 0 and r.state.converged, (n, r.state.relerr)
        b = of.gen_bumps_grid(_desc(n), seed, nb)
        rel = np.linalg.norm(b - oracle.poisson_apply(n, [1.0 / n[0]] * len(n), [(D, D)] * len(n), outs[k])) \
            / np.linalg.norm(b)
        print(f"{n}: iters {r.state.niter} refine {r.state.nrefine} engine relres {r.state.relerr:.2e} "
              f"oracle relres {rel:.2e}", flush=True)
        assert rel <= 1.01e-8
    alone, _ = _serve(ctx, [LARGE[3]])
    np.testing.assert_array_equal(alone[0], outs[3])
