"""Host-side exhaustive checks of the device kernels' index math (no GPU)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def _build_run(name, tmp_path):
    exe = str(tmp_path / name)
    src = os.path.join(HERE, "cpp_host", name + ".cpp")
    inc = os.path.join(HERE, "..", "opflow_amd", "csrc")
    subprocess.run(["g++", "-O1", "-std=c++17", "-fsanitize=address,undefined", "-I" + inc, src, "-o", exe],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_ghost_index_math(tmp_path):
    assert "passed" in _build_run("index_check", tmp_path)


def test_fastdiv_matches_hardware_divide(tmp_path):
    assert "passed" in _build_run("fastdiv_check", tmp_path)
