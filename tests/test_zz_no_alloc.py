"""The allocation-free FTCS entry point (opf_ftcs_noalloc), called through the
C ABI, against opf_ftcs on the same inputs: bit-identical results on two calls
in a row on one device (the second on a smaller field, so the scratch holds the
larger call's leftovers), 2-D and 3-D, FP64 and FP32, launch plans that end on
either buffer; oversized fields are refused before any work."""
import ctypes as C

import numpy as np
import pytest

import opflow_amd as of
from opflow_amd import _lib as L
import oracle
from helpers import build_field, random_storage, SEED

pytestmark = pytest.mark.gpu
DIRC, NEUM, PER = 1, 2, 3

# (dim, bc per dim, dtype, [(cells per side, steps) for call 1, call 2])
CASES = [
    (2, [(DIRC, DIRC), (NEUM, DIRC)], "f64", [(200, 37), (48, 13)]),
    (2, [(PER, PER), (DIRC, NEUM)], "f32", [(160, 21), (40, 9)]),
    (3, [(PER, PER)] * 3, "f64", [(64, 14), (24, 13)]),
    (3, [(PER, PER)] * 3, "f32", [(48, 7), (20, 5)]),
    (3, [(DIRC, DIRC), (NEUM, NEUM), (PER, PER)], "f64", [(40, 6), (18, 5)]),
]


def _field(ctx, dim, n, bc, dtype, seed):
    b = of.MeshBuilder(ctx).newMesh(*([n + 1] * dim))
    for d in range(dim):
        b.setMeshOfDim(d, 0.0, 1.0)
    mesh = b.build()
    om = oracle.Mesh([n + 1] * dim)
    bcv = [(0.25 if t == DIRC else -0.5 if t == NEUM else 0.0 for t in sides) for sides in bc]
    bcv = [tuple(v) for v in bcv]
    u = build_field(ctx, mesh, dim, 1, bc, bcv, dtype=dtype)
    v = build_field(ctx, mesh, dim, 1, bc, bcv, dtype=dtype)
    init = oracle.update_padding(u.desc, om, random_storage(u, seed))
    u.upload(init)
    v.upload(init)
    return u, v, om, init


@pytest.mark.parametrize("case", range(len(CASES)))
def test_noalloc_equals_ftcs_two_calls(ctx, case):
    lib = L.load()
    dim, bc, dtype, calls = CASES[case]
    for call, (n, steps) in enumerate(calls):
        u, v, om, init = _field(ctx, dim, n, bc, dtype, SEED + 40 + 7 * case + call)
        assert lib.opf_field_nbytes(v.h) <= L.OPF_FTCS_NOALLOC_MAX_BYTES
        h = 1.0 / n
        coef = 0.2 * h * h if dim == 2 else 0.15 * h * h
        print(f"case {case} call {call}: dim {dim} n {n} {dtype} steps {steps} plan {v.ftcs_plan(steps)}",
              flush=True)
        u.ftcs(coef, steps)
        rc = lib.opf_ftcs_noalloc(v.h, C.c_double(coef), steps)
        assert rc == L.OPF_OK, lib.opf_last_error().decode()
        got, ref = v.numpy(), u.numpy()
        np.testing.assert_array_equal(got, ref)
        if dtype == "f64" and call == 1:
            np.testing.assert_array_equal(got, oracle.ftcs(v.desc, om, init, coef, steps))
        u.close()
        v.close()


def test_noalloc_refuses_oversized_field(ctx):
    lib = L.load()
    n = 3000  # 3002^2 doubles = 72.1 MB > 64 MiB
    mesh = of.MeshBuilder(ctx).newMesh(n + 1, n + 1).setMeshOfDim(0, 0., 1.).setMeshOfDim(1, 0., 1.).build()
    u = build_field(ctx, mesh, 2, 1, [(DIRC, DIRC), (DIRC, DIRC)])
    assert lib.opf_field_nbytes(u.h) > L.OPF_FTCS_NOALLOC_MAX_BYTES
    before = u.numpy()
    rc = lib.opf_ftcs_noalloc(u.h, C.c_double(1e-8), 3)
    assert rc == L.OPF_ERR_NOMEM
    assert b"exceeds" in lib.opf_last_error()
    np.testing.assert_array_equal(u.numpy(), before)
    u.close()
