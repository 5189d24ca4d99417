"""Shared test wiring: builds matching product fields and oracle meshes."""
import ctypes as C

import numpy as np

import opflow_amd as of
from opflow_amd import _lib as L
import oracle

SEED = 0xC0FFEE


def desc(dim, n_nodes, loc, bc, bcv=None, ext=(1, 1), padding=0, dtype=L.OPF_F64):
    """Host-only finalized FieldDesc (no GPU). bc: per dim (start, end) types."""
    d = L.FieldDesc()
    d.dim, d.dtype, d.padding = dim, dtype, padding
    for q in range(dim):
        d.loc[q] = loc[q] if isinstance(loc, (list, tuple)) else loc
        for s in range(2):
            d.bc_type[q][s] = bc[q][s]
            d.bc_value[q][s] = bcv[q][s] if bcv is not None else 0.0
            d.ext[q][s] = ext[s] if isinstance(ext[0], int) else ext[q][s]
    start = (C.c_int * 3)(0, 0, 0)
    nn = (C.c_int * 3)(*(list(n_nodes) + [1] * (3 - dim)))
    L.check(L.load().opf_field_desc_finalize(C.byref(d), start, nn))
    return d


def build_field(ctx, mesh, dim, loc, bc, bcv=None, ext=(1, 1), padding=0, dtype="f64", name="u"):
    b = of.ExprBuilder().setName(name).setMesh(mesh).setDType(dtype).setPadding(padding)
    b.setLoc(list(loc) if isinstance(loc, (list, tuple)) else loc)
    for q in range(dim):
        for s in range(2):
            b.setBC(q, s, bc[q][s], bcv[q][s] if bcv is not None else 0.0)
            b.setExt(q, s, ext[s] if isinstance(ext[0], int) else ext[q][s])
    return b.build()


def random_storage(field_or_desc, seed, lo=-1.0, hi=1.0):
    d = field_or_desc.desc if hasattr(field_or_desc, "desc") else field_or_desc
    n = d.extent[0] * d.extent[1] * d.extent[2]
    a = of.gen_uniform(seed, n, lo, hi)
    shape = tuple(d.extent[q] for q in reversed(range(d.dim)))
    return a.reshape(shape)


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))
