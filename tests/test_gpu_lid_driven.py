"""Lid-driven cavity (MAC grid, explicit advection/diffusion + pressure
projection through Solve with a pinned Neumann Poisson): the mixed-precision
run agrees with the FP64 run, the projected velocity is divergence-free to the
solver tolerance, and the flow develops the expected primary vortex."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "opflow_amd", "bin")


def _run(n, steps, mixed, out):
    r = subprocess.run([os.path.join(BIN, "lid_driven2d"), str(n), str(steps), str(int(mixed)), out],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_lid_driven_mixed_vs_fp64(tmp_path):
    n, steps = 128, 100
    a = _run(n, steps, False, str(tmp_path / "a.bin"))
    b = _run(n, steps, True, str(tmp_path / "b.bin"))
    ua = np.fromfile(tmp_path / "a.bin")
    ub = np.fromfile(tmp_path / "b.bin")
    assert np.isfinite(ua).all() and np.isfinite(ub).all()
    rel = np.max(np.abs(ua - ub)) / np.max(np.abs(ua))
    print("max |u_mixed - u_fp64| / max|u| =", rel)
    assert rel < 1e-6
    assert a["worst_relres"] <= 1e-8 and b["worst_relres"] <= 1e-8
    # projection: discrete divergence small compared with |u|/h
    assert a["max_div"] < 1e-5 * n
    # the lid drags the fluid: u below the lid is positive, the centre return flow negative
    assert a["u_center"] < 0
