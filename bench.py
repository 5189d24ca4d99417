"""Benchmark of the MI355X engine on OpFlow's explicit hot path.

Headline workload (one GPU): 3-D FTCS heat sweep, 512^3 cells FP64,
periodic, u0 ~ U[-1,1] i.i.d. from a seed, dt = 0.2 h^2, alpha = 1 -- one
"step" is one FTCS time step over the whole grid (a launch fuses up to three).
value = cell-updates/s over the timed steps (inputs resident in HBM).

Also reported (extra): the FP32 variant of the same sweep, the 2-D 256^2
Dirichlet FTCS config (1000 steps), and the implicit-solver numbers when the
multigrid solver is built.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-updates/s & %HBM peak; Poisson solves/s + p99 latency at rel-res 1e-8"
HBM_PEAK_GBS = 8000.0


class HipEvents:
    """hipEventRecord on the library's stream (not torch's)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]

    def create(self):
        e = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def ftcs_field(of, ctx, n, dim, dtype, bc_type, seed, init="uniform"):
    nodes = [n + 1] * dim
    mb = of.MeshBuilder(ctx).newMesh(*nodes)
    for d in range(dim):
        mb.setMeshOfDim(d, 0.0, 1.0)
    mesh = mb.build()
    b = of.ExprBuilder().setName("u").setMesh(mesh).setLoc(of.LocOnMesh.Center).setExt(1).setDType(dtype)
    for d in range(dim):
        b.setBC(d, of.DimPos.start, bc_type, 0.0).setBC(d, of.DimPos.end, bc_type, 0.0)
    u = b.build()
    if init == "uniform":
        import numpy as np
        cnt = 1
        for e in u.shape:
            cnt *= e
        u.upload(of.gen_uniform(seed, cnt, -1.0, 1.0).reshape(u.shape).astype(u.dtype))
    else:
        u.upload(of.gen_bumps(u, seed, 4).astype(u.dtype))
    u.updatePadding()
    return mesh, u


CONFIG_STEPS = 200  # the 3-D configuration's step count (FP32-vs-FP64 accuracy is compared there)
SETTLE_STEPS = 320  # untimed steps before any timed region: after only 64, 200 timed steps at 512^3 ran 2.5 % (FP64) and 17 % (FP32) below steady state (clock ramp)


def time_ftcs(of, ctx, ev, u, coef, steps, warmup):
    """HIP-event time of `steps` FTCS steps on the library's stream (every
    launch in the region is an FTCS kernel launch), after max(warmup, SETTLE_STEPS)
    untimed steps."""
    u.ftcs(coef, max(warmup, SETTLE_STEPS))
    ctx.sync()
    e0, e1 = ev.create(), ev.create()
    t0 = time.perf_counter()
    ev.record(e0, ctx.stream)
    u.ftcs(coef, steps)
    ev.record(e1, ctx.stream)
    ctx.sync()
    wall = time.perf_counter() - t0
    ms = ev.elapsed_ms(e0, e1)
    return ms, wall


def generic_sweeps(of, ctx, ev, n):
    """The generic `u = expr` path (opf_sweep) on the headline shape: the FTCS update
    written as an expression (3-D 7-point, FP64 and FP32) and a WENO53 advection step
    (2-D 2048^2, 3 ghost layers), with the compiled-expression engine (default) and the
    stack-program interpreter. One step per launch; the first (compiling) sweep is untimed."""
    import numpy as np
    D2, W = of.D2SecondOrderCentered, of.D1WENO53Upwind

    def timed(fn, reps):
        fn()
        ctx.sync()
        e0, e1 = ev.create(), ev.create()
        ev.record(e0, ctx.stream)
        for _ in range(reps):
            fn()
        ev.record(e1, ctx.stream)
        return ev.elapsed_ms(e0, e1) / reps

    out = {}
    for dtype in ("f64", "f32"):
        _, u = ftcs_field(of, ctx, n, 3, dtype, of.BCType.Periodic, 0xC0FFEE + 2)
        expr = u + (0.2 / n / n) * (of.d2x(D2, u) + of.d2y(D2, u) + of.d2z(D2, u))
        for mode, tag in ((1, "compiled"), (0, "interpreter")):
            ctx.set_jit(mode)
            ms = timed(lambda: u.assign(expr), 20 if mode else 3)
            s = 16.0 if dtype == "f64" else 8.0
            out[f"sweep_expr_ftcs3d_{n}_{dtype}_{tag}"] = {
                "Mcell_updates_per_s": n ** 3 / (ms / 1e3) / 1e6, "ms_per_step": ms,
                "alg_GBps": s * n ** 3 / (ms / 1e3) / 1e9}
        u.close()
    ctx.set_jit(1)
    m = 2048
    mesh = of.MeshBuilder(ctx).newMesh(m + 1, m + 1).setMeshOfDim(0, 0., 1.).setMeshOfDim(1, 0., 1.).build()
    b = of.ExprBuilder().setName("u").setMesh(mesh).setLoc(of.LocOnMesh.Center).setExt(3)
    for d in range(2):
        b.setBC(d, of.DimPos.start, of.BCType.Periodic).setBC(d, of.DimPos.end, of.BCType.Periodic)
    u = b.build()
    u.upload(of.gen_uniform(0xC0FFEE + 7, int(np.prod(u.shape)), -1.0, 1.0).reshape(u.shape))
    u.updatePadding()
    v = u.clone("v")
    expr = u - 1e-4 * (of.dx(W, u) + of.dy(W, u))
    ms = timed(lambda: v.assign(expr), 20)
    out[f"sweep_expr_weno53_2d_{m}_f64_compiled"] = {"Mcell_updates_per_s": m * m / (ms / 1e3) / 1e6,
                                                      "ms_per_step": ms, "alg_GBps": 16.0 * m * m / (ms / 1e3) / 1e9}
    u.close()
    v.close()
    return out


def cpp_programs(n):
    """The FTCS time loops written as OpFlow programs (tests/cpp/ftcs2d_example.cpp, ftcs3d_example.cpp)
    on the C++ front-end, which defers the repeated assignment into fused launches. 3-D: after the
    first assignment and the same untimed clock-settling steps as the headline, the time from the
    next assignment to the device finishing the last; 2-D: until the host read of the result."""
    import subprocess
    bind = os.path.join(ROOT, "opflow_amd", "bin")
    out = {}
    for dt_ in ("f64", "f32"):
        r = subprocess.run([os.path.join(bind, "ftcs3d_example"), str(n), str(CONFIG_STEPS), dt_, "",
                            str(SETTLE_STEPS)], capture_output=True, text=True, timeout=300)
        if r.returncode == 0:
            j = json.loads(r.stdout.strip().splitlines()[-1])
            out[f"cpp_ftcs3d_{n}_{dt_}_{CONFIG_STEPS}steps"] = {"Mcell_updates_per_s": j["Mcell_updates_per_s"],
                                                                 "s_loop": j["s_loop"]}
        else:
            out[f"cpp_ftcs3d_{n}_{dt_}_{CONFIG_STEPS}steps"] = {"error": r.stderr[-300:]}
    r = subprocess.run([os.path.join(bind, "ftcs2d_example"), "", "256", "1000"], capture_output=True, text=True,
                       timeout=300)
    if r.returncode == 0:
        j = json.loads([x for x in r.stdout.splitlines() if x.startswith("{")][0])
        out["cpp_ftcs2d_256_fp64_1000steps"] = {
            "ms_loop_and_read": j["s_loop_and_read"] * 1e3, "ms_first_step": j["s_first_step"] * 1e3,
            "ms_steps_2_to_1000_and_read": j["s_steps_2_to_n_and_read"] * 1e3,
            "Mcell_updates_per_s_steps_2_to_1000": 256 * 256 * 999 / j["s_steps_2_to_n_and_read"] / 1e6,
            "note": "the first step also allocates the second time level and loads the kernels (first use)"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=24,
                    help="steps of each timed run of the 3-D FTCS CPU baseline (median of 5 runs)")
    ap.add_argument("--dtype", choices=["f64", "f32"], default="f64",
                    help="precision of the headline sweep (f64 is the configuration's headline)")
    ap.add_argument("--probe-copy", action="store_true",
                    help="only run 5 device copies of one 512^3 FP64 field (2 x nbytes of HBM traffic each): "
                         "the known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: the engine is single-GPU by design "
              "(a slab decomposition along z with ghost exchange is the planned sharding, "
              "not implemented)", file=sys.stderr)
        sys.exit(2)

    import opflow_amd as of
    from opflow_amd import _lib as L
    ctx = of.Context(0)
    ev = HipEvents()
    n = args.n
    if args.probe_copy:
        _, a = ftcs_field(of, ctx, n, 3, "f64", of.BCType.Periodic, 1)
        _, b = ftcs_field(of, ctx, n, 3, "f64", of.BCType.Periodic, 2)
        for _ in range(5):
            b.copy_from(a)
        ctx.sync()
        nbytes = 8 * a.shape[0] * a.shape[1] * a.shape[2]
        print(json.dumps({"probe": "copy", "bytes_read_per_copy": nbytes, "bytes_written_per_copy": nbytes}))
        return
    h = 1.0 / n
    coef = 0.2 * h * h  # dt = 0.2 h^2, alpha = 1

    # --- headline: 3-D FTCS 512^3 FP64 periodic
    mesh, u = ftcs_field(of, ctx, n, 3, args.dtype, of.BCType.Periodic, 0xC0FFEE + 2)
    ms, wall = time_ftcs(of, ctx, ev, u, coef, args.steps, args.warmup)
    cells = n ** 3
    value = cells * args.steps / (ms / 1e3) / 1e6
    launches, steps_per_launch = u.ftcs_plan(args.steps)
    per_launch_s = ms / 1e3 / launches
    # algorithmic bytes of one launch: read u once and write it once (FP64),
    # whatever number of steps the launch fuses
    alg_bytes = (16.0 if args.dtype == "f64" else 8.0) * cells
    achieved = alg_bytes / per_launch_s / 1e9
    u.close()

    extra = {}
    if not args.no_extra and args.dtype == "f64":
        mesh32, u32 = ftcs_field(of, ctx, n, 3, "f32", of.BCType.Periodic, 0xC0FFEE + 2)
        ms32, _ = time_ftcs(of, ctx, ev, u32, coef, args.steps, args.warmup)
        v32 = cells * args.steps / (ms32 / 1e3) / 1e6
        l32, k32 = u32.ftcs_plan(args.steps)
        extra["ftcs3d_512_fp32"] = {"Mcell_updates_per_s": v32, "ms_per_step": ms32 / args.steps,
                                    "steps_per_launch": k32,
                                    "alg_GBps_per_launch": 8.0 * cells / (ms32 / 1e3 / l32) / 1e9}
        extra["ftcs3d_512_fp32"]["hbm_frac"] = extra["ftcs3d_512_fp32"]["alg_GBps_per_launch"] / HBM_PEAK_GBS
        u32.close()
        # the same sweep one step per launch (no temporal blocking): the streaming kernel's
        # share of HBM peak, 16 B per cell-update (FP64) / 8 B (FP32)
        for dt_, bpc in (("f64", 16.0), ("f32", 8.0)):
            _, us = ftcs_field(of, ctx, n, 3, dt_, of.BCType.Periodic, 0xC0FFEE + 2)
            us.ftcs_blocked(coef, SETTLE_STEPS, 1)
            ctx.sync()
            e0, e1 = ev.create(), ev.create()
            ev.record(e0, ctx.stream)
            us.ftcs_blocked(coef, 40, 1)
            ev.record(e1, ctx.stream)
            ctx.sync()
            ms1 = ev.elapsed_ms(e0, e1) / 40
            gbs = bpc * cells / (ms1 / 1e3) / 1e9
            extra[f"ftcs3d_512_{dt_}_one_step_per_launch"] = {
                "Mcell_updates_per_s": cells / (ms1 / 1e3) / 1e6, "ms_per_step": ms1, "alg_GBps": gbs,
                "hbm_frac": gbs / HBM_PEAK_GBS}
            us.close()
        # accuracy of the configuration: FP32 against FP64 after the configuration's
        # 200 steps (whatever --steps is) from the same input (the FP64 GPU sweep
        # equals the oracle bit for bit), relative to max|u_FP64|; untimed, fresh fields
        import numpy as np
        _, a64 = ftcs_field(of, ctx, n, 3, "f64", of.BCType.Periodic, 0xC0FFEE + 2)
        a64.ftcs(coef, CONFIG_STEPS)
        r64 = a64.numpy()[1:-1, 1:-1, 1:-1]
        a64.close()
        _, a32 = ftcs_field(of, ctx, n, 3, "f32", of.BCType.Periodic, 0xC0FFEE + 2)
        a32.ftcs(coef, CONFIG_STEPS)
        d32 = a32.numpy()[1:-1, 1:-1, 1:-1].astype(np.float64) - r64
        a32.close()
        scale = float(np.max(np.abs(r64)))
        extra["ftcs3d_512_fp32"].update({
            "steps_compared": CONFIG_STEPS, "max_abs_u_fp64": scale,
            "max_err_rel": float(np.max(np.abs(d32))) / scale,
            "l2_err_rel": float(np.linalg.norm(d32) / np.linalg.norm(r64))})
        del d32, r64
        # stable control (dt = h^2/8, below the 7-point limit h^2/6), labelled as such:
        # the same comparison where the solution decays instead of blowing up
        coef_s = h * h / 8.0
        outs = []
        for dt_ in ("f64", "f32"):
            _, us = ftcs_field(of, ctx, n, 3, dt_, of.BCType.Periodic, 0xC0FFEE + 2)
            us.ftcs(coef_s, CONFIG_STEPS)
            outs.append(us.numpy()[1:-1, 1:-1, 1:-1].astype(np.float64))
            us.close()
        d32 = outs[1] - outs[0]
        scale = float(np.max(np.abs(outs[0])))
        extra["ftcs3d_512_stable_control_dt_h2_over_8"] = {
            "steps_compared": CONFIG_STEPS, "max_abs_u_fp64": scale,
            "fp32_max_err_rel": float(np.max(np.abs(d32))) / scale,
            "fp32_l2_err_rel": float(np.linalg.norm(d32) / np.linalg.norm(outs[0]))}
        del d32, outs
        n2 = 256
        h2 = 1.0 / n2
        mesh2, u2 = ftcs_field(of, ctx, n2, 2, "f64", of.BCType.Dirc, 0xC0FFEE + 1, init="bumps")
        init2 = (u2.numpy(), L.FieldDesc.from_buffer_copy(u2.desc))
        ms2, _ = time_ftcs(of, ctx, ev, u2, 0.2 * h2 * h2, 1000, 50)
        extra["ftcs2d_256_fp64_1000steps"] = {"Mcell_updates_per_s": n2 * n2 * 1000 / (ms2 / 1e3) / 1e6,
                                              "ms_total": ms2}
        u2.close()
        extra.update(generic_sweeps(of, ctx, ev, n))
        extra.update(cpp_programs(n))
        # 3-D FTCS with Dirichlet sides (blocked, ghosts rebuilt inside each launch)
        _, ud = ftcs_field(of, ctx, n, 3, "f64", of.BCType.Dirc, 0xC0FFEE + 2)
        msd, _ = time_ftcs(of, ctx, ev, ud, 0.125 * h * h, 60, 10)
        ld, kd = ud.ftcs_plan(60)
        extra["ftcs3d_512_fp64_dirichlet_60steps"] = {
            "Mcell_updates_per_s": cells * 60 / (msd / 1e3) / 1e6, "ms_per_step": msd / 60, "steps_per_launch": kd,
            "alg_GBps_per_launch": 16.0 * cells / (msd / 1e3 / ld) / 1e9}
        ud.close()
        from bench_solver import solver_extra
        extra.update(solver_extra(of, ctx, ev, cpu=not args.no_cpu))

    cpu = None
    if not args.no_cpu:
        import bench_cpu
        import oracle
        from tests.helpers import desc as mkdesc
        d3 = mkdesc(3, [n + 1] * 3, 1, [(3, 3)] * 3, ext=(1, 1))
        om3 = oracle.Mesh([n + 1] * 3)
        cnt = d3.extent[0] * d3.extent[1] * d3.extent[2]
        init3 = of.gen_uniform(0xC0FFEE + 2, cnt, -1.0, 1.0).reshape(d3.extent[2], d3.extent[1], d3.extent[0])
        init3 = oracle.update_padding(d3, om3, init3)
        b3 = bench_cpu.ftcs3d_512(init3, d3, om3, coef, n=n, steps_all=args.cpu_steps)
        del init3
        a = b3["all_threads"]
        cpu = {"value": a["Mcell_updates_per_s"], "unit": "Mcell-updates/s", "cores": a["threads"], "kind": "port",
               "sample": f"oracle FTCS (OpFlow FieldAssigner sequence, C/OpenMP) on the same 512^3 FP64 periodic "
                         f"input, {args.cpu_steps} steps per run, median of {a['runs']} runs after a warm-up run "
                         f"({a['s_per_run']:.2f} s per run) on {a['threads']} threads of {cpu_model()}"}
        if not args.no_extra and args.dtype == "f64":
            cb = extra.setdefault("cpu_baselines", {})
            cb["cpu_model"] = cpu_model()
            cb["ftcs3d_512_fp64"] = b3
            d2 = init2[1]
            cb["ftcs2d_256_fp64_1000steps"] = bench_cpu.ftcs2d_256(init2[0], d2, oracle.Mesh([n2 + 1] * 2),
                                                                   0.2 * h2 * h2)
            cb["lid_driven_1024sq_per_step"] = bench_cpu.lid_driven(of, ctx)
            from bench_solver import _requests
            cb["served_poisson_one_at_a_time"] = bench_cpu.served(of, _requests(512, 0xC0FFEE + 5)[:16])

    traffic = None
    tf = os.path.join(ROOT, "profiles", "ftcs3d_512_fp64_traffic.json")
    if os.path.exists(tf) and args.dtype == "f64" and n == 512:
        try:
            traffic = json.load(open(tf)).get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None

    line = {
        "metric": METRIC, "value": value, "unit": "Mcell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "untimed_steps_before_timing": max(args.warmup, SETTLE_STEPS),
        "ms_per_step": ms / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": args.dtype, "data": "synthetic (i.i.d. U[-1,1] from SplitMix64 seed 0xC0FFEE+2)",
        "config": {"workload": f"ftcs3d 512^3 cells {'fp64' if args.dtype == 'f64' else 'fp32'} periodic dt=0.2h^2",
                   "cells": cells, "grid": [n, n, n], "bc": "periodic"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "step_equivalent_GBps": alg_bytes * args.steps / (ms / 1e3) / 1e9,
                     "step_equivalent_frac": alg_bytes * args.steps / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                     "traffic_GBps": traffic / per_launch_s / 1e9 if traffic else None,
                     "traffic_frac": traffic / per_launch_s / 1e9 / HBM_PEAK_GBS if traffic else None,
                     "kernel": "ftcs3d_tb_kernel" if steps_per_launch > 1 else "ftcs3d_lds_kernel",
                     "steps_per_launch": steps_per_launch, "launches": launches},
        "cpu_baseline": cpu,
        "wall_s_timed_region": wall,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
