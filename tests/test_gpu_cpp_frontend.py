"""The C++ OpFlow front-end (opflow_amd/include/OpFlow) driving the engine:
the FTCS example and the reference's equation tests restated in C++."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from oracle import cart
from helpers import desc

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "opflow_amd", "bin")


def _run(args, timeout=240):
    exe = os.path.join(BIN, args[0])
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built: run __graft_entry__.build()")
    r = subprocess.run([exe] + args[1:], capture_output=True, text=True, timeout=timeout)
    print(r.stdout, r.stderr, flush=True)
    return r


def test_ftcs2d_example_bitexact(tmp_path):
    n, steps = 256, 1000
    out = str(tmp_path / "u.bin")
    r = _run(["ftcs2d_example", out, str(n), str(steps), str(tmp_path) + "/"])
    assert r.returncode == 0
    raw = np.fromfile(out, dtype=np.float64)
    m = (n + 2) * (n + 2)
    u, u2, init = raw[:m].reshape(n + 2, n + 2), raw[m:2 * m].reshape(n + 2, n + 2), raw[2 * m:].reshape(n + 2, n + 2)
    c = cart.read(str(tmp_path / "u_0.cart"))  # RawBinaryOStream output of the same field
    assert c["name"] == "u" and c["time"] == steps * (0.2 / (n * n)) and c["local_range"] == [(0, n), (0, n)]
    np.testing.assert_array_equal(c["values"], u[1:-1, 1:-1].reshape(-1))
    d = desc(2, [n + 1, n + 1], 1, [(1, 1), (1, 1)], ext=(1, 1))
    om = oracle.Mesh([n + 1, n + 1])
    init = oracle.update_padding(d, om, init.copy())
    coef = 0.2 / (n * n) * 1.0
    np.testing.assert_array_equal(u, oracle.ftcs(d, om, init, coef, steps))
    np.testing.assert_array_equal(u2, oracle.ftcs(d, om, init, coef, 5))


def test_cpp_equation_tests():
    r = _run(["eqn_tests"])
    assert r.returncode == 0, r.stdout
    assert r.stdout.count("PASS") == 15 and "FAIL" not in r.stdout


@pytest.mark.parametrize("n,steps,dtype", [(40, 14, "f64"), (24, 70, "f64"), (32, 9, "f32")])
def test_ftcs3d_example_deferred_steps(tmp_path, n, steps, dtype):
    """The unchanged OpFlow time loop `u = u + dt*alpha*(d2x(u)+d2y(u)+d2z(u))` on a periodic
    3-D field: the front-end runs the first step at once and defers the rest into fused
    launches (flushed every 60 steps and at the sync); the result equals the oracle's FTCS
    sequence bit for bit (FP64) or to FP32 rounding."""
    out = str(tmp_path / "u.bin")
    r = _run(["ftcs3d_example", str(n), str(steps), dtype, out])
    assert r.returncode == 0
    d = desc(3, [n + 1] * 3, 1, [(3, 3)] * 3, ext=(1, 1))
    om = oracle.Mesh([n + 1] * 3)
    import opflow_amd as of
    cnt = d.extent[0] * d.extent[1] * d.extent[2]
    init = of.gen_uniform(0xC0FFEE + 2, cnt, -1.0, 1.0).reshape(d.extent[2], d.extent[1], d.extent[0])
    init = oracle.update_padding(d, om, init)
    h = 1.0 / n
    ref = oracle.ftcs(d, om, init, 0.2 * h * h * 1.0, steps)  # dt * alpha as the program forms it
    got = np.fromfile(out, dtype=np.float64 if dtype == "f64" else np.float32).reshape(ref.shape)
    if dtype == "f64":
        np.testing.assert_array_equal(got, ref)
    else:
        assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))


def test_ftcs3d_example_settle_steps(tmp_path):
    """The untimed settle assignments (argv[5]) are ordinary time steps: 9 steps after 65
    settle steps (a deferred flush falls inside them) leave the field the oracle's 74-step
    FTCS sequence gives, bit for bit, and the JSON line reports them."""
    import json
    n, steps, settle = 24, 9, 65
    out = str(tmp_path / "u.bin")
    r = _run(["ftcs3d_example", str(n), str(steps), "f64", out, str(settle)])
    assert r.returncode == 0
    j = json.loads(r.stdout.strip().splitlines()[-1])
    assert j["settle"] == settle and j["timed_steps"] == steps - 1
    d = desc(3, [n + 1] * 3, 1, [(3, 3)] * 3, ext=(1, 1))
    om = oracle.Mesh([n + 1] * 3)
    import opflow_amd as of
    cnt = d.extent[0] * d.extent[1] * d.extent[2]
    init = of.gen_uniform(0xC0FFEE + 2, cnt, -1.0, 1.0).reshape(d.extent[2], d.extent[1], d.extent[0])
    init = oracle.update_padding(d, om, init)
    h = 1.0 / n
    ref = oracle.ftcs(d, om, init, 0.2 * h * h * 1.0, steps + settle)
    np.testing.assert_array_equal(np.fromfile(out, dtype=np.float64).reshape(ref.shape), ref)


def test_conv1d_example_against_oracle(tmp_path):
    """examples/CONV1D/CONV1D.cpp on the front-end: 100 steps of first-order upwind advection of
    a square pulse, and the conservative Superbee-limited variant, equal to the oracle's
    evaluation of the same expressions bit for bit; the limited scheme conserves mass and
    stays within the initial bounds."""
    import json
    import opflow_amd as of
    from oracle import expr as E
    out = str(tmp_path / "c.bin")
    r = _run(["conv1d_example", out, "100"])
    assert r.returncode == 0
    info = json.loads(r.stdout.strip().splitlines()[-1])
    du = desc(1, [101], 0, [(1, 1)], ext=(0, 0))
    dw = desc(1, [101], 0, [(1, 1)], ext=(2, 2))
    om = oracle.Mesh([101])
    raw = np.fromfile(out)
    nu_, nw_ = du.extent[0], dw.extent[0]
    u_got, w_got = raw[:nu_], raw[nu_:nu_ + nw_]

    def init(d):
        x = om.x(0, np.arange(d.offset[0], d.offset[0] + d.extent[0]))
        return oracle.update_padding(d, om, np.where((0.2 <= x) & (x <= 0.4), 1.0, 0.0))
    dt, c = 0.5e-2, 1.0
    # u - dt*c*dx<D1FirstOrderBiasedDownwind>(u): FIELD u, CONST, D1DOWN, MUL, SUB
    nodes_u = [(E.FIELD, 0, 0, 0, 0.0), (E.CONST, 0, 0, 0, dt * c), (E.D1DOWN, 0, 0, 0, 0.0),
               (E.MUL, 1, 2, 0, 0.0), (E.SUB, 0, 3, 0, 0.0)]
    # w - dt*dx<D1FirstOrderCentered>(c * intp<Superbee, Cor2Cen>(c, w))
    code = float(E.LIM_SUPERBEE + 32 * 2)
    nodes_w = [(E.FIELD, 0, 0, 0, 0.0), (E.CONST, 0, 0, 0, c), (E.FLUXLIM, 1, 0, 0, code), (E.MUL, 1, 2, 0, 0.0),
               (E.D1C, 3, 0, 0, 0.0), (E.CONST, 0, 0, 0, dt), (E.MUL, 5, 4, 0, 0.0), (E.SUB, 0, 6, 0, 0.0)]
    u, w = init(du), init(dw)
    m0 = w[dw.assign_start[0] - dw.offset[0]:dw.assign_end[0] - dw.offset[0]].sum()
    for _ in range(100):
        u = E.assign(nodes_u, 4, [(u, du)], om, u, du, 0, update_padding=lambda a: oracle.update_padding(du, om, a))
        w = E.assign(nodes_w, 7, [(w, dw)], om, w, dw, 0, update_padding=lambda a: oracle.update_padding(dw, om, a))
    np.testing.assert_array_equal(u_got, u)
    np.testing.assert_array_equal(w_got, w)
    assert abs(info["w_sum"] - m0) <= 1e-12 * m0
    assert info["w_min"] >= -1e-15 and info["w_max"] <= 1.0 + 1e-15
    assert info["u_min"] >= 0.0 and info["u_max"] <= 1.0
