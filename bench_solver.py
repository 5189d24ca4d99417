"""Implicit-solver and serving legs of bench.py (extra keys of the JSON line).

cfg3: backward-Euler/Poisson solves at 2048^2 and 256^3 (Dirichlet, bumps +
1e-3 noise RHS) to relative residual 1e-10 -- FP64 GMG-FCG vs FP32-inner/FP64-IR.
cfg5: served Poisson solves of mixed sizes at tolerance 1e-8 -- solves/s closed
loop and p50/p99 latency open loop, next to the plain FP64 solver one at a time.
"""
import time

import numpy as np


def _solve_timed(of, ctx, ev, desc, b, precision, tol, reps=3):
    g = of.PoissonGMG(ctx, desc, precision)
    n = b.size
    bb, xx = of.DeviceBuffer(ctx, n), of.DeviceBuffer(ctx, n)
    bb.upload(b)
    st = g.solve(bb.ptr, xx.ptr, tol=tol)  # warm-up (also builds nothing new)
    times = []
    for _ in range(reps):
        xx.upload(np.zeros(n))
        ctx.sync()
        t0 = time.perf_counter()
        st = g.solve(bb.ptr, xx.ptr, tol=tol)
        ctx.sync()
        times.append(time.perf_counter() - t0)
    x = xx.numpy()
    g.close(); bb.close(); xx.close()
    return x, st, float(np.median(times))


def cfg3(of, ctx, ev):
    out = {}
    for n in ([2048, 2048], [256, 256, 256]):
        dim = len(n)
        desc = of.poisson_desc(dim, n, [1.0 / n[0]] * dim, [(1, 1)] * dim)
        b = of.gen_bumps_grid(desc, 0xC0FFEE + 3, 4, 1e-3)
        x64, s64, t64 = _solve_timed(of, ctx, ev, desc, b, "fp64", 1e-10)
        x32, s32, t32 = _solve_timed(of, ctx, ev, desc, b, "mixed", 1e-10)
        err = float(np.max(np.abs(x32 - x64)) / np.max(np.abs(x64)))
        key = "x".join(map(str, n))
        out[f"poisson_{key}_relres1e-10"] = {
            "fp64": {"s_per_solve": t64, "iters": s64.niter, "relres": s64.relerr},
            "mixed": {"s_per_solve": t32, "iters": s32.niter, "refinements": s32.nrefine, "relres": s32.relerr,
                      "max_err_vs_fp64": err},
            "mixed_speedup": t64 / t32,
        }
    return out


def _requests(count, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(count):
        if rng.random() < 0.5:
            s = int(rng.choice([64, 128, 256, 512, 1024, 2048]))
            n = [s, s]
        else:
            s = int(rng.choice([32, 64, 128, 256]))
            n = [s, s, s]
        reqs.append((n, int(rng.integers(1, 1 << 30)), int(rng.integers(1, 9))))
    return reqs


def _desc(of, n):
    dim = len(n)
    return of.poisson_desc(dim, n, [1.0 / n[0]] * dim, [(1, 1)] * dim)


def _fp64_one_at_a_time(of, ctx, reqs, tol, arrivals=None):
    """Plain FP64 GPU solver serving one request at a time (no batching, solver set up per shape once)."""
    cache = {}
    lat, t_start = [], time.perf_counter()
    clock = 0.0
    for k, (n, seed, nb) in enumerate(reqs):
        key = tuple(n)
        if key not in cache:
            g = of.PoissonGMG(ctx, _desc(of, n), "fp64")
            N = int(np.prod(n))
            cache[key] = (g, of.DeviceBuffer(ctx, N), of.DeviceBuffer(ctx, N))
        g, bb, xx = cache[key]
        b = of.gen_bumps_grid(_desc(of, n), seed, nb)
        if arrivals is not None:
            clock = max(clock, arrivals[k])
        t0 = time.perf_counter()
        bb.upload(b)
        xx.upload(np.zeros(b.size))
        g.solve(bb.ptr, xx.ptr, tol=tol)
        ctx.sync()
        dt = time.perf_counter() - t0
        if arrivals is not None:
            clock += dt
            lat.append(clock - arrivals[k])
    total = time.perf_counter() - t_start
    for g, bb, xx in cache.values():
        g.close(); bb.close(); xx.close()
    return total, lat


def _serve_run(of, ctx, reqs, tol, arrivals=None):
    srv = of.Server(ctx, 64 << 30, max_inflight=16, precision="mixed", batched=True, tol=tol)
    t0 = time.perf_counter()
    pending = list(range(len(reqs)))
    done, lat = 0, []
    base = srv.clock()
    while done < len(reqs):
        now = srv.clock() - base
        while pending and (arrivals is None or arrivals[pending[0]] <= now):
            k = pending.pop(0)
            n, seed, nb = reqs[k]
            t_arr = -1.0 if arrivals is None else base + arrivals[k]
            srv.submit(_desc(of, n), seed, nb, t_submit=t_arr)
        if srv.pending() == (0, 0) and pending:
            time.sleep(max(0.0, min(0.002, arrivals[pending[0]] - now)))
            continue
        srv.step()
        for r in srv.poll():
            done += 1
            lat.append(r.t_done - r.t_submit)
            assert r.state.converged
    total = time.perf_counter() - t0
    srv.close()
    return total, lat


def cfg5(of, ctx, ev, count=48):
    tol = 1e-8
    reqs = _requests(count, 0xC0FFEE + 5)
    # warm-up: first touch of every shape (hierarchies, graphs)
    _serve_run(of, ctx, reqs[:8], tol)
    t_srv, _ = _serve_run(of, ctx, reqs, tol)
    t_64, _ = _fp64_one_at_a_time(of, ctx, reqs, tol)
    # open loop: Poisson arrivals at a rate that keeps the FP64 one-at-a-time server at ~85% utilisation
    lam = 0.85 * count / t_64
    rng = np.random.default_rng(0xC0FFEE + 55)
    arr = np.cumsum(rng.exponential(1.0 / lam, count))
    _, lat_srv = _serve_run(of, ctx, reqs, tol, arrivals=arr)
    _, lat_64 = _fp64_one_at_a_time(of, ctx, reqs, tol, arrivals=arr)
    pct = lambda v, q: float(np.percentile(np.asarray(v), q))
    return {"served_poisson_mixed_sizes_relres1e-8": {
        "requests": count, "arrival_rate_per_s": lam,
        "served_mixed_batched": {"solves_per_s": count / t_srv, "p50_s": pct(lat_srv, 50), "p99_s": pct(lat_srv, 99)},
        "fp64_one_at_a_time": {"solves_per_s": count / t_64, "p50_s": pct(lat_64, 50), "p99_s": pct(lat_64, 99)},
        "throughput_ratio": t_64 / t_srv}}


def solver_extra(of, ctx, ev):
    out = {}
    out.update(cfg3(of, ctx, ev))
    out.update(cfg5(of, ctx, ev))
    return out
