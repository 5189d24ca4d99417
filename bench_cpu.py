"""CPU baselines of bench.py: the oracle's restatement of OpFlow's CPU path,
timed on the host cores of the machine that runs the benchmark, for every
configuration. Reported beside the GPU figures, never a target.

Each entry gives the rate at all host threads and at one thread, the median
of several timed runs after one untimed warm-up run, and the sample it was
measured on (a bounded share of the configuration where a full run would take
minutes on the CPU). The restatements:
  * FTCS (2-D and 3-D): oracle FTCS in C/OpenMP, OpFlow's FieldAssigner
    sequence per step (copy to a temporary, evaluate, updatePadding, copy back;
    FieldAssigner.hpp:26-86);
  * Poisson / pressure solves: the oracle's C/OpenMP flexible PCG with a
    V(1,1) weighted-Jacobi multigrid cycle (the PFMG defaults, pfmg.c:25-39);
  * lid-driven momentum and projection sweeps: the oracle's numpy evaluation
    of the same expression trees (vectorised, one thread).
"""
import statistics
import time

import numpy as np


def _median(fn, runs):
    fn()  # warm-up
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts), len(ts)


def _at_threads(nthreads, fn):
    import oracle
    oracle.set_threads(nthreads)
    try:
        return fn(), oracle.num_threads()
    finally:
        oracle.set_threads(0)


def ftcs2d_256(init, desc, om, coef, steps=1000, runs=5):
    """2-D FTCS 256^2 Dirichlet, 1000 steps (the whole configuration per run)."""
    import oracle
    out = {}
    for tag, nt in (("all_threads", 0), ("one_thread", 1)):
        (t, k), cores = _at_threads(nt, lambda: _median(lambda: oracle.ftcs(desc, om, init, coef, steps, nt), runs))
        out[tag] = {"Mcell_updates_per_s": 256 * 256 * steps / t / 1e6, "s_per_run": t, "runs": k,
                    "threads": cores, "sample": f"{steps} steps, the whole configuration"}
    return out


def ftcs3d_512(init, desc, om, coef, n=512, steps_all=24, steps_one=1, runs=5):
    """3-D FTCS n^3 periodic FP64: `steps_all` steps per run at all threads, `steps_one` at one."""
    import oracle
    out = {}
    for tag, nt, st in (("all_threads", 0, steps_all), ("one_thread", 1, steps_one)):
        (t, k), cores = _at_threads(nt, lambda: _median(lambda: oracle.ftcs(desc, om, init, coef, st, nt), runs))
        out[tag] = {"Mcell_updates_per_s": n ** 3 * st / t / 1e6, "s_per_run": t, "runs": k, "threads": cores,
                    "sample": f"{st} of the 200 steps per run from the configuration's input"}
    return out


def poisson_one(of, n, b, tol=1e-10):
    """One implicit problem (grid n, right-hand side b) to 1e-10: median of 5 runs (3 for 3-D) at
    all threads; at one thread, the same right-hand-side recipe on a grid of half the side.
    Returns the timings and the all-threads solution."""
    import oracle
    dim = len(n)
    out, x_full = {}, None
    for tag, nt in (("all_threads", 0), ("one_thread", 1)):
        m = n if nt == 0 else [k // 2 for k in n]
        bb = b if nt == 0 else of.gen_bumps_grid(of.poisson_desc(dim, m, [1.0 / m[0]] * dim, [(1, 1)] * dim),
                                                 0xC0FFEE + 3, 4, 1e-3)
        runs = (5 if dim == 2 else 3) if nt == 0 else 1
        res = {}

        def run():
            res["x"], res["it"], res["rel"] = oracle.pcg_mg(m, [1.0 / m[0]] * dim, [(1, 1)] * dim, bb, tol=tol,
                                                            maxit=400)
        (t, k), cores = _at_threads(nt, lambda: _median(run, runs))
        out[tag] = {"s_per_solve": t, "solves_per_s": 1.0 / t, "iters": res["it"], "relres": res["rel"], "runs": k,
                    "threads": cores, "grid": m,
                    "sample": "the configuration's problem" if m == n else f"{'x'.join(map(str, m))} sample"}
        if nt == 0:
            x_full = res["x"]
    return out, x_full


def lid_driven(of, ctx, n=1024, steps=5, steps_one=2):
    """Lid-driven cavity 1024^2 MAC: one time step = the momentum sweeps (du, dv, u += du, v += dv),
    the pressure solve (Neumann, mean-zero right-hand side, warm-started from the previous
    pressure, to 1e-8) and the projection sweeps, from the state after one warm-up step."""
    import oracle
    from oracle import expr as oexpr
    DIRC, NEUM = 1, 2
    h = 1.0 / n
    nu = 1e-3
    dt = min(0.25 * h * h / nu, 0.5 * h)
    mesh = of.MeshBuilder(ctx).newMesh(n + 1, n + 1).setMeshOfDim(0, 0., 1.).setMeshOfDim(1, 0., 1.).build()
    om = oracle.Mesh([n + 1, n + 1])

    def fld(name, loc, bc, bcv):
        b = of.ExprBuilder().setName(name).setMesh(mesh).setLoc(loc).setExt(1)
        for d in range(2):
            for s in range(2):
                b.setBC(d, s, bc[d][s], bcv[d][s])
        return b.build()
    U, C_ = of.LocOnMesh.Corner, of.LocOnMesh.Center
    u = fld("u", [U, C_], [(DIRC, DIRC)] * 2, [(0, 0), (0, 1.0)])
    du = fld("du", [U, C_], [(DIRC, DIRC)] * 2, [(0, 0), (0, 1.0)])
    v = fld("v", [C_, U], [(DIRC, DIRC)] * 2, [(0, 0), (0, 0)])
    dv = fld("dv", [C_, U], [(DIRC, DIRC)] * 2, [(0, 0), (0, 0)])
    p = fld("p", [C_, C_], [(NEUM, NEUM)] * 2, [(0, 0), (0, 0)])
    D1, D2 = of.D1FirstOrderCentered, of.D2SecondOrderCentered
    I_c2n, I_n2c = of.d1IntpCenterToCorner, of.d1IntpCornerToCenter
    exprs = {
        "du": dt * (nu * (of.d2x(D2, u) + of.d2y(D2, u)) - (of.dx(D1, I_n2c(0, u) * I_n2c(0, u)) +
                                                            of.dy(D1, I_c2n(1, u) * I_c2n(0, v)))),
        "dv": dt * (nu * (of.d2x(D2, v) + of.d2y(D2, v)) - (of.dx(D1, I_c2n(1, u) * I_c2n(0, v)) +
                                                            of.dy(D1, I_n2c(1, v) * I_n2c(1, v)))),
        "u1": u + du, "v1": v + dv,
        "div": (of.dx(D1, u) + of.dy(D1, v)) / dt,
        "u2": u - dt * of.dx(D1, p), "v2": v - dt * of.dy(D1, p)}
    flat = {k: e.flatten() for k, e in exprs.items()}
    target = {"du": du, "dv": dv, "u1": u, "v1": v, "u2": u, "v2": v}
    st = {"u": np.zeros(u.shape), "v": np.zeros(v.shape), "du": np.zeros(du.shape), "dv": np.zeros(dv.shape),
          "p": np.zeros(p.shape)}
    for k, f in (("u", u), ("v", v), ("du", du), ("dv", dv), ("p", p)):
        st[k] = oracle.update_padding(f.desc, om, st[k])

    def assign(key, out_name):
        nodes, flds, root = flat[key]
        out = target[key]
        st[out_name] = oexpr.assign(nodes, root, [(st[f.name], f.desc) for f in flds], om, st[out_name], out.desc,
                                    0, update_padding=lambda r: oracle.update_padding(out.desc, om, r))

    iters = []

    def step():
        assign("du", "du")
        assign("dv", "dv")
        assign("u1", "u")
        assign("v1", "v")
        nodes, flds, root = flat["div"]
        rhs = oexpr.evaluate(nodes, root, [(st[f.name], f.desc) for f in flds], om, [0, 0], [n, n]).ravel()
        rhs = rhs - rhs.mean()
        x0 = np.ascontiguousarray(st["p"][1:-1, 1:-1]).ravel()
        x, it, _ = oracle.pcg_mg([n, n], [h, h], [(NEUM, NEUM)] * 2, -rhs, x0=x0, tol=1e-8)
        iters.append(it)
        pp = st["p"].copy()
        pp[1:-1, 1:-1] = x.reshape(n, n)
        st["p"] = oracle.update_padding(p.desc, om, pp)
        assign("u2", "u")
        assign("v2", "v")

    out = {}
    for tag, nt, k in (("all_threads", 0, steps), ("one_thread", 1, steps_one)):
        iters.clear()
        (t, runs), cores = _at_threads(nt, lambda: _median(step, k))
        out[tag] = {"s_per_step": t, "runs": runs, "threads": cores, "pcg_iters_per_step": float(np.mean(iters)),
                    "sample": f"{runs} time steps after a warm-up step, from rest; sweeps in numpy (one "
                              f"thread), pressure solve on {cores} thread(s)"}
    for f in (u, v, du, dv, p):
        f.close()
    return out


def served(of, reqs, tol=1e-8, one_thread_max_cells=1 << 18):
    """Served mix: the requests solved one at a time (no batching), all threads; the one-thread
    figure covers the requests of at most 2^18 cells of the same sample."""
    import oracle
    out = {}
    for tag, nt in (("all_threads", 0), ("one_thread", 1)):
        sample = reqs if nt == 0 else [r for r in reqs if int(np.prod(r[0])) <= one_thread_max_cells]
        rhs = []
        for n, seed, nb in sample:
            dim = len(n)
            rhs.append(of.gen_bumps_grid(of.poisson_desc(dim, n, [1.0 / n[0]] * dim, [(1, 1)] * dim), seed, nb))

        def run():
            for (n, _, _), b in zip(sample, rhs):
                dim = len(n)
                oracle.pcg_mg(n, [1.0 / n[0]] * dim, [(1, 1)] * dim, b, tol=tol, maxit=400)
        oracle.set_threads(nt)
        try:
            t0 = time.perf_counter()
            run()
            t = time.perf_counter() - t0
            cores = oracle.num_threads()
        finally:
            oracle.set_threads(0)
        out[tag] = {"solves_per_s": len(sample) / t, "requests": len(sample), "s_total": t, "threads": cores,
                    "sample": f"{len(sample)} requests of the seeded stream, one at a time, one pass" +
                              ("" if nt == 0 else f" (those of <= {one_thread_max_cells} cells)")}
    return out
