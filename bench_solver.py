"""Implicit-solver and serving legs of bench.py (extra keys of the JSON line).

cfg3: backward-Euler/Poisson solves at 2048^2 and 256^3 (Dirichlet, bumps +
1e-3 noise RHS) to relative residual 1e-10 -- FP64 GMG-FCG vs FP32-inner/FP64-IR
(and, with cpu=True, the CPU restatement of the reference's solver).
cfg4: lid-driven cavity 1024^2, 500 steps, mixed vs FP64 pressure solves.
cfg5: served Poisson solves of mixed sizes at tolerance 1e-8 -- solves/s closed
loop and p50/p99 latency open loop, next to the plain FP64 solver one at a time.
eqn_sets: SolveEqns on ten equal-operator equations (batched) vs ten Solve()
calls, and a coupled two-equation set.
eqn_setup: per-solve setup of a non-static handler (pads and multigrid levels
built on the device) on a variable-coefficient operator at 1024^2 and 256^3.
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


def cfg3(of, ctx, ev, cpu=False):
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
        N = int(np.prod(n))
        # algorithmic bytes of one PCG iteration at maximal fusion (1 pre + 1 post smoothing):
        # 24.3 s N in 2-D, 22.4 s N in 3-D (s = bytes per value); mixed: FP32 iterations plus
        # 6 FP64 vectors per refinement (true residual + update)
        c = 24.3 if dim == 2 else 22.4
        b64 = c * 8 * N * s64.niter
        b32 = c * 4 * N * s32.niter + 6 * 8 * N * s32.nrefine
        ent = out[f"poisson_{key}_relres1e-10"]
        for tag, byt, t in (("fp64", b64, t64), ("mixed", b32, t32)):
            ent[tag]["roofline"] = {"bound": "hbm", "achieved": byt / t / 1e9, "peak": 8000.0, "unit": "GB/s",
                                    "frac": byt / t / 1e9 / 8000.0, "alg_bytes_per_solve": byt}
        if cpu:
            # reported CPU baseline: the oracle's C/OpenMP restatement of the reference's CPU
            # solver (flexible PCG + V(1,1) weighted-Jacobi multigrid, PFMG defaults), median
            # of several runs at all threads, and a smaller sample at one thread
            import bench_cpu
            ent["cpu_baseline"], xc = bench_cpu.poisson_one(of, n, b)
            ent["cpu_baseline"].update({"kind": "port",
                                        "max_err_vs_gpu_fp64": float(np.max(np.abs(xc - x64)) / np.max(np.abs(x64)))})
    return out


def cfg4(of, ctx, ev, n=1024, steps=500):
    """Lid-driven cavity 1024^2 MAC, Re 1000, 500 steps (the C++ OpFlow program
    opflow_amd/examples/lid_driven2d.cpp): mixed-precision pressure solves vs FP64."""
    import json
    import os
    import subprocess
    import tempfile
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "opflow_amd", "bin", "lid_driven2d")
    if not os.path.exists(exe):
        return {"lid_driven_1024sq_500steps": {"skipped": "opflow_amd/bin/lid_driven2d not built"}}
    res = {}
    # an untimed 5-step run of each precision first: it compiles the program's expression
    # kernels into the on-disk cache, which a fresh box starts without (the FP64 run, going
    # first, had otherwise paid for compiles inside its timed steps: 1.16-1.19 vs 0.88 ms)
    for mixed in (0, 1):
        subprocess.run([exe, str(n), "5", str(mixed), ""], capture_output=True, text=True, timeout=600)
    with tempfile.TemporaryDirectory() as td:
        for mixed in (0, 1):
            path = os.path.join(td, f"u{mixed}.bin")
            r = subprocess.run([exe, str(n), str(steps), str(mixed), path], capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                return {"lid_driven_1024sq_500steps": {"error": r.stderr[-400:]}}
            res[mixed] = (json.loads(r.stdout.strip().splitlines()[-1]), np.fromfile(path))
    (a, ua), (b, ub) = res[0], res[1]
    return {"lid_driven_1024sq_500steps": {
        "fp64": {k: a[k] for k in ("s_per_step", "s_first_step", "poisson_iters", "worst_relres", "max_div", "dt")},
        "mixed": {k: b[k] for k in ("s_per_step", "s_first_step", "poisson_iters", "worst_relres", "max_div")},
        "timing": "s_per_step over steps 2..500 (after an untimed 5-step run per precision that fills the "
                  "compiled-kernel cache); the first step also builds the hierarchy (s_first_step)",
        "velocity_max_rel_diff_mixed_vs_fp64": float(np.max(np.abs(ua - ub)) / np.max(np.abs(ua)))}}


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


class _Fp64Server:
    """Plain FP64 GPU solver serving one request at a time: one resident
    solver per shape, RHS generated on the device from the request's seed."""

    def __init__(self, of, ctx):
        self.of, self.ctx, self.cache = of, ctx, {}

    def solve(self, n, seed, nb, tol):
        of = self.of
        key = tuple(n)
        if key not in self.cache:
            N = int(np.prod(n))
            self.cache[key] = (of.PoissonGMG(self.ctx, _desc(of, n), "fp64"), of.DeviceBuffer(self.ctx, N),
                               of.DeviceBuffer(self.ctx, N))
        g, bb, xx = self.cache[key]
        of.gen_bumps_grid_dev(self.ctx, _desc(of, n), seed, nb, 0.0, bb)
        xx.zero()
        st = g.solve(bb.ptr, xx.ptr, tol=tol)
        assert st.converged
        return st

    def run(self, reqs, tol, arrivals=None):
        lat, clock = [], 0.0
        t_start = time.perf_counter()
        for k, (n, seed, nb) in enumerate(reqs):
            if arrivals is not None:
                clock = max(clock, arrivals[k])
            t0 = time.perf_counter()
            self.solve(n, seed, nb, tol)
            dt = time.perf_counter() - t0
            if arrivals is not None:
                clock += dt
                lat.append(clock - arrivals[k])
        return time.perf_counter() - t_start, lat

    def close(self):
        for g, bb, xx in self.cache.values():
            g.close(); bb.close(); xx.close()


def _alg_bytes(n, st):
    """Algorithmic bytes of one mixed-precision solve: per PCG iteration 24.3 s N (2-D) or
    22.4 s N (3-D) at maximal fusion with s = 4 B, plus 6 FP64 vectors per refinement."""
    N = int(np.prod(n))
    c = 24.3 if len(n) == 2 else 22.4
    return c * 4 * N * st.niter + 6 * 8 * N * st.nrefine


def _serve_run(of, srv, reqs, tol, arrivals=None):
    """Submit reqs (all at once, or at their arrival times on the server clock) and run rounds until done."""
    t0 = time.perf_counter()
    pending = list(range(len(reqs)))
    done, lat, byt, ids = 0, [], 0.0, {}
    base = srv.clock()
    while done < len(reqs):
        now = srv.clock() - base
        while pending and (arrivals is None or arrivals[pending[0]] <= now):
            k = pending.pop(0)
            n, seed, nb = reqs[k]
            t_arr = -1.0 if arrivals is None else base + arrivals[k]
            ids[srv.submit(_desc(of, n), seed, nb, t_submit=t_arr)] = k
        if srv.pending() == (0, 0) and pending:
            time.sleep(max(0.0, min(0.002, arrivals[pending[0]] - now)))
            continue
        srv.step()
        for r in srv.poll():
            done += 1
            lat.append(r.t_done - r.t_submit)
            assert r.state.converged
            byt += _alg_bytes(reqs[ids[r.id]][0], r.state)
    return time.perf_counter() - t0, lat, byt


def cfg5(of, ctx, ev, count=512):
    tol = 1e-8
    reqs = _requests(count, 0xC0FFEE + 5)
    shapes = sorted({tuple(n) for n, _, _ in reqs})
    # admission by memory footprint (64 GiB budget); the in-flight cap is above the request count
    srv = of.Server(ctx, 64 << 30, max_inflight=64, precision="mixed", batched=True, tol=tol)
    base = _Fp64Server(of, ctx)
    # steady state: both servers are warmed with one pass of the same request list, so the
    # state they keep resident (groups / per-shape solvers) exists before timing (nothing
    # else carries over: every request regenerates its RHS and starts from x = 0)
    _serve_run(of, srv, reqs, tol)
    base.run(reqs, tol)
    t_srv, _, byt_srv = _serve_run(of, srv, reqs, tol)
    t_64, _ = base.run(reqs, tol)
    # open loop: Poisson arrivals at a rate that keeps the FP64 one-at-a-time server at ~85% utilisation
    lam = 0.85 * count / t_64
    rng = np.random.default_rng(0xC0FFEE + 55)
    arr = np.cumsum(rng.exponential(1.0 / lam, count))
    _, lat_srv, _ = _serve_run(of, srv, reqs, tol, arrivals=arr)
    _, lat_64 = base.run(reqs, tol, arrivals=arr)
    srv.close()
    base.close()
    pct = lambda v, q: float(np.percentile(np.asarray(v), q))
    return {"served_poisson_mixed_sizes_relres1e-8": {
        "requests": count, "shapes": [list(n) for n in shapes], "arrival_rate_per_s": lam,
        "server": {"mem_budget_GiB": 64, "max_inflight": 64, "warm_up": "one untimed pass of the same request list (both servers)"},
        "rhs": "seeded bumps generated on the device at admission (both servers)",
        "served_mixed_batched": {"solves_per_s": count / t_srv, "p50_s": pct(lat_srv, 50), "p99_s": pct(lat_srv, 99),
                                 "roofline": {"bound": "hbm", "achieved": byt_srv / t_srv / 1e9, "peak": 8000.0,
                                              "unit": "GB/s", "frac": byt_srv / t_srv / 1e9 / 8000.0,
                                              "note": "closed loop; algorithmic bytes of each request's FP32 "
                                                      "iterations and FP64 refinements"}},
        "fp64_one_at_a_time": {"solves_per_s": count / t_64, "p50_s": pct(lat_64, 50), "p99_s": pct(lat_64, 99)},
        "throughput_ratio": t_64 / t_srv}}


def eqn_sets(of, ctx, n=256, neq=10, reps=3):
    """SolveEqns (UnifiedSolve.hpp:69-100) on ten copies of a Dirichlet Poisson equation
    on n^2 cells with different boundary values: the set (one batched hierarchy, ten
    slots) next to ten single-equation Solve() calls; each call includes the device
    pad evaluation, the hierarchy build, the solve and the write-back, as the
    reference's SolveEqns / Solve do per call. Plus a coupled two-equation set."""
    D2 = of.D2SecondOrderCentered
    mesh = of.MeshBuilder(ctx).newMesh(n + 1, n + 1).build()

    def field(name, g):
        b = of.ExprBuilder().setMesh(mesh).setName(name).setLoc([of.LocOnMesh.Center] * 2).setExt(1)
        for d in range(2):
            b.setBC(d, of.DimPos.start, of.BCType.Dirc, g).setBC(d, of.DimPos.end, of.BCType.Dirc, g)
        return b.build()
    ps = [field(f"p{i}", 0.1 * i) for i in range(neq)]

    def lap(e):
        return of.d2x(D2, e) + of.d2y(D2, e)

    def getter(i):
        return lambda *e: of.Equation(1.0, lap(e[i]))
    eqs = [getter(i) for i in range(neq)]
    params = [of.IJSolverParams(tol=1e-10)] + [of.IJSolverParams() for _ in range(neq - 1)]
    ts, st = [], None
    for r in range(reps + 1):
        for q in ps:
            q.assign(0.0)
        ctx.sync()
        t0 = time.perf_counter()
        st = of.SolveEqns(*eqs, *ps, None, params)
        ctx.sync()
        ts.append(time.perf_counter() - t0)
    t_set = min(ts[1:])
    ts1 = []
    for r in range(reps + 1):
        for q in ps:
            q.assign(0.0)
        ctx.sync()
        t0 = time.perf_counter()
        for q in ps:
            of.Solve(lambda e: 1.0 == lap(e), q, of.StructSolverParams(tol=1e-10))
        ctx.sync()
        ts1.append(time.perf_counter() - t0)
    t_one = min(ts1[1:])
    # coupled: diffusion-reaction of two species exchanging mass
    u, v = field("u", 0.0), field("v", 1.0)
    ceqs = [lambda a, b: of.Equation(1.0, lap(a) - 50.0 * a + 20.0 * b),
            lambda a, b: of.Equation(0.0, lap(b) - 40.0 * b + 30.0 * a)]
    tc = []
    for r in range(reps + 1):
        u.assign(0.0)
        v.assign(0.0)
        ctx.sync()
        t0 = time.perf_counter()
        cst = of.SolveEqns(*ceqs, u, v, None, [of.IJSolverParams(tol=1e-10), of.IJSolverParams()])
        ctx.sync()
        tc.append(time.perf_counter() - t0)
    return {f"solve_eqns_{neq}x_poisson_{n}sq_relres1e-10": {
        "set_s_per_call": t_set, "set_iters": st.niter, "set_relres": st.relerr,
        "single_solves_s_total": t_one, "speedup": t_one / t_set,
        "note": "per call: device pad evaluation, hierarchy build, solve, write-back (both sides)"},
        f"solve_eqns_coupled_2x{n}sq_relres1e-10": {"s_per_call": min(tc[1:]), "bicgstab_iters": cst.niter,
                                                     "relres": cst.relerr}}


def eqn_setup(of, ctx, reps=3):
    """Per-solve setup of a non-static handler (HYPREEqnSolveHandler::solve with staticMat
    false: generateAb on every solve, HYPREEqnSolveHandler.hpp:125-166,190-209) on the
    DircEqnTest operator d/dx(dp/dx / r) + d/dy(dp/dy / r) (r = 1 + 9 x y, a variable
    coefficient) at 1024^2 and 256^3: a non-static solve (pads, levels built on the
    device, solve) beside a static one (right-hand side only, same solve); the difference
    is the setup. FP64, tol 1e-8, from the previous solution (the solve does the same
    iterations both ways)."""
    D1, C2N = of.D1FirstOrderCentered, of.d1IntpCenterToCorner
    out = {}
    for n in ([1024, 1024], [256, 256, 256]):
        dim = len(n)
        mb = of.MeshBuilder(ctx).newMesh(*[k + 1 for k in n])
        for d in range(dim):
            mb.setMeshOfDim(d, 0.0, 1.0)
        mesh = mb.build()

        def field(name, bc, g):
            b = of.ExprBuilder().setMesh(mesh).setName(name).setLoc([of.LocOnMesh.Center] * dim).setExt(1)
            for d in range(dim):
                b.setBC(d, of.DimPos.start, bc, g).setBC(d, of.DimPos.end, bc, g)
            return b.build()
        p, r, f = field("p", of.BCType.Dirc, 0.0), field("r", of.BCType.Neum, 0.0), field("f", of.BCType.Dirc, 0.0)
        r.initBy(lambda x: 1.0 + 9.0 * x[0] * x[1])
        f.initBy(lambda x: np.sin(3 * x[0]) * np.cos(2 * x[1]) + x[0])
        ops = [of.dx, of.dy, of.dz][:dim]

        def eqn(e):
            lhs = None
            for d, op in enumerate(ops):
                t = op(D1, op(D1, e) / C2N(d, r))
                lhs = t if lhs is None else lhs + t
            return f == lhs
        res = {}
        for static in (False, True):
            h = of.makeEqnSolveHandler(eqn, p, of.PrecondStructSolver(of.StructSolverParams(tol=1e-8,
                                                                                              staticMat=static)))
            p.assign(0.0)
            h.solve()  # first solve: full evaluation and build either way
            ts = []
            for _ in range(reps):
                ctx.sync()
                t0 = time.perf_counter()
                st = h.solve()
                ctx.sync()
                ts.append(time.perf_counter() - t0)
            res[static] = (float(np.median(ts)), st)
            h.close()
        key = "x".join(map(str, n))
        out[f"eqn_setup_varcoef_{key}"] = {
            "nonstatic_s_per_solve": res[False][0], "static_s_per_solve": res[True][0],
            "setup_s": res[False][0] - res[True][0], "iters": res[False][1].niter, "relres": res[False][1].relerr,
            "note": "non-static: pads + device level construction + solve; static: right-hand side + solve"}
        for q in (p, r, f):
            q.close()
    return out


def solver_extra(of, ctx, ev, cpu=False):
    """Every leg above; a leg that raises is reported as its error (the bench line still prints)."""
    out = {}
    legs = (("cfg3", lambda: cfg3(of, ctx, ev, cpu)), ("cfg4", lambda: cfg4(of, ctx, ev)),
            ("cfg5", lambda: cfg5(of, ctx, ev)), ("eqn_sets", lambda: eqn_sets(of, ctx)),
            ("eqn_setup", lambda: eqn_setup(of, ctx)))
    for name, leg in legs:
        try:
            out.update(leg())
        except Exception as e:  # noqa: BLE001
            out[f"{name}_error"] = repr(e)[:400]
    return out
