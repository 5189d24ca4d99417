"""Builds the C++ front-end programs (examples and C++ parity tests) that
include opflow_amd/include/OpFlow and link libopflow_hip.so."""
import glob
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BIN = os.path.join(HERE, "bin")


def build_cpp():
    srcs = sorted(glob.glob(os.path.join(ROOT, "tests", "cpp", "*.cpp")) +
                  glob.glob(os.path.join(HERE, "examples", "*.cpp")))
    if not srcs:
        return []
    os.makedirs(BIN, exist_ok=True)
    out = []
    for s in srcs:
        exe = os.path.join(BIN, os.path.basename(s)[:-4])
        cmd = ["g++", "-O2", "-std=c++20", "-ffp-contract=off", "-I" + os.path.join(HERE, "include"),
               "-I" + os.path.join(ROOT, "include"), s, "-o", exe, "-L" + os.path.join(HERE, "lib"),
               "-lopflow_hip", "-Wl,-rpath,$ORIGIN/../lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("C++ build failed: " + " ".join(cmd) + "\n" + r.stderr)
        out.append(exe)
    return out
