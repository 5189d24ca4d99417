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
    assert r.stdout.count("PASS") == 10 and "FAIL" not in r.stdout
