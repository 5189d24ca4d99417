"""OpFlow's raw binary field file (`.cart`) in numpy -- TEST INFRASTRUCTURE ONLY.

Restates the byte layout RawBinaryOStream::operator<< writes and
RawBinaryIStream::operator>> reads (Utils/Writers/RawBinaryStream.hpp:98-224):
int name length, the name's bytes, int dim, int rank count, double time, per
dim int (mesh start, end), per dim the node coordinates over the mesh range,
per dim int (start, end) of the accessible range, the same for the local
range, then the local range's values x-fastest in the field's element type.
Native byte order (little-endian here), no padding between items.
"""
import struct

import numpy as np


def write(path, name, time, mesh_range, mesh_x, access_range, local_range, values, dtype=np.float64):
    """values: array over the local range in (z, y, x) order (or x-fastest flat)."""
    dim = len(mesh_range)
    nb = name.encode()
    out = [struct.pack("<i", len(nb)), nb, struct.pack("<ii", dim, 1), struct.pack("<d", float(time))]
    for s, e in mesh_range:
        out.append(struct.pack("<ii", s, e))
    for d in range(dim):
        out.append(np.asarray(mesh_x[d], dtype="<f8").tobytes())
    for rng in (access_range, local_range):
        for s, e in rng:
            out.append(struct.pack("<ii", s, e))
    out.append(np.ascontiguousarray(values, dtype=np.dtype(dtype).newbyteorder("<")).reshape(-1).tobytes())
    with open(path, "wb") as f:
        f.write(b"".join(out))


def read(path, dtype=np.float64):
    """Returns a dict with name, dim, nproc, time, mesh_range, mesh_x, access_range,
    local_range and values (flat, x-fastest)."""
    raw = open(path, "rb").read()
    p = 0

    def take(fmt):
        nonlocal p
        v = struct.unpack_from(fmt, raw, p)
        p += struct.calcsize(fmt)
        return v

    (n,) = take("<i")
    name = raw[p:p + n].decode()
    p += n
    dim, nproc = take("<ii")
    (time,) = take("<d")
    mesh_range = [take("<ii") for _ in range(dim)]
    mesh_x = []
    for s, e in mesh_range:
        mesh_x.append(np.frombuffer(raw, dtype="<f8", count=e - s, offset=p).copy())
        p += 8 * (e - s)
    access_range = [take("<ii") for _ in range(dim)]
    local_range = [take("<ii") for _ in range(dim)]
    cnt = 1
    for s, e in local_range:
        cnt *= e - s
    dt = np.dtype(dtype).newbyteorder("<")
    values = np.frombuffer(raw, dtype=dt, count=cnt, offset=p).copy()
    p += dt.itemsize * cnt
    if p != len(raw):
        raise ValueError(f"{path}: {len(raw) - p} trailing bytes")
    return dict(name=name, dim=dim, nproc=nproc, time=time, mesh_range=mesh_range, mesh_x=mesh_x,
                access_range=access_range, local_range=local_range, values=values)
