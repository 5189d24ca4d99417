"""Field streams on the GPU: RawBinaryOStream / RawBinaryIStream `.cart` files
(Utils/Writers/RawBinaryStream.hpp:84-224) written from and read into device
fields, checked byte for byte against the oracle's restatement of the layout
(oracle/cart.py), including files larger than one staging chunk."""
import os
import time

import numpy as np
import pytest

import opflow_amd as of
import oracle
from oracle import cart
from helpers import SEED, build_field, random_storage

pytestmark = pytest.mark.gpu
D_, N_, P_ = 1, 2, 3


def _mesh(ctx, nodes):
    return of.MeshBuilder(ctx).newMesh(*nodes).build() if len(nodes) > 1 else None


def _local(f, arr):
    lr = f.localRange
    sl = tuple(slice(lr[q][0] - f.desc.offset[q], lr[q][1] - f.desc.offset[q]) for q in reversed(range(f.desc.dim)))
    return arr[sl]


def _expected_bytes(tmp, f, mesh, t, dtype):
    path = os.path.join(tmp, "expected.cart")
    cart.write(path, f.name, t, mesh.getRange(), [mesh.xs[d][mesh.pad:mesh.pad + mesh.dims[d]]
                                                 for d in range(mesh.dim)],
               f.accessibleRange, f.localRange, _local(f, f.numpy()), dtype)
    return open(path, "rb").read()


@pytest.mark.parametrize("dim,dtype", [(2, "f64"), (3, "f32")])
def test_write_matches_oracle_layout(ctx, tmp_path, dim, dtype):
    nodes = [33, 17] if dim == 2 else [9, 12, 7]
    mesh = of.MeshBuilder(ctx).newMesh(*nodes).setMeshOfDim(0, 0., 2.).build()
    bc = [(D_, N_), (P_, P_), (N_, D_)][:dim]
    u = build_field(ctx, mesh, dim, 1, bc, bcv=[(0.5, -1.0), (0, 0), (0.25, 0.0)][:dim], dtype=dtype, name="u")
    u.upload(random_storage(u, SEED + 31))
    u.updatePadding()
    os_ = of.Utils.RawBinaryOStream(str(tmp_path) + "/")
    os_ << of.Utils.TimeStamp(0.125) << u << of.Utils.TimeStamp(0.25) << u
    npdt = np.float64 if dtype == "f64" else np.float32
    want = _expected_bytes(str(tmp_path), u, mesh, 0.125, npdt)
    assert open(tmp_path / "u_0.cart", "rb").read() == want
    r = cart.read(str(tmp_path / "u_1.cart"), npdt)
    assert r["time"] == 0.25 and r["name"] == "u" and r["nproc"] == 1


def test_read_fills_local_range_and_updates_padding(ctx, tmp_path):
    mesh = of.MeshBuilder(ctx).newMesh(21, 13).setMeshOfDim(1, -1., 1.).build()
    om = oracle.Mesh([21, 13], [(0., 1.), (-1., 1.)])
    p = build_field(ctx, mesh, 2, 1, [(D_, N_), (N_, D_)], bcv=[(1.5, 0.0), (0.2, -0.7)], name="p")
    vals = of.gen_uniform(SEED + 32, 20 * 12, -1.0, 1.0).reshape(12, 20)
    cart.write(str(tmp_path / "p_4.cart"), "p", 7.5, mesh.getRange(),
               [mesh.xs[d][mesh.pad:mesh.pad + mesh.dims[d]] for d in range(2)], p.accessibleRange, p.localRange,
               vals)
    is_ = of.Utils.RawBinaryIStream(str(tmp_path) + "/")
    is_.setCounterTo(4)
    is_ >> p
    assert is_.time == 7.5
    ref = np.zeros(p.shape)
    _local(p, ref)[...] = vals
    ref = oracle.update_padding(p.desc, om, ref)
    np.testing.assert_array_equal(p.numpy(), ref)


def test_read_rejects_layout_mismatch(ctx, tmp_path):
    mesh = of.MeshBuilder(ctx).newMesh(9, 9).build()
    u = build_field(ctx, mesh, 2, 1, [(D_, D_), (D_, D_)], name="u")
    of.Utils.RawBinaryOStream(str(tmp_path) + "/") << u
    mesh2 = of.MeshBuilder(ctx).newMesh(9, 11).build()
    v = build_field(ctx, mesh2, 2, 1, [(D_, D_), (D_, D_)], name="u")
    with pytest.raises(of._lib.OpfError, match="mismatch"):
        of.Utils.RawBinaryIStream(str(tmp_path) + "/") >> v


def test_round_trip_larger_than_a_chunk(ctx, tmp_path):
    # 257^3 nodes -> 256^3 = 16.8 M cells: three staging chunks each way
    mesh = of.MeshBuilder(ctx).newMesh(257, 257, 257).build()
    u = build_field(ctx, mesh, 3, 1, [(P_, P_)] * 3, name="big")
    v = build_field(ctx, mesh, 3, 1, [(P_, P_)] * 3, name="big")
    u.upload(random_storage(u, SEED + 33))
    u.updatePadding()
    ctx.sync()
    t0 = time.perf_counter()
    of.Utils.RawBinaryOStream(str(tmp_path) + "/") << u
    t1 = time.perf_counter()
    of.Utils.RawBinaryIStream(str(tmp_path) + "/") >> v
    t2 = time.perf_counter()
    mb = 256 ** 3 * 8 / 1e6
    print(f"256^3 FP64 .cart: write {mb / (t1 - t0):.0f} MB/s, read {mb / (t2 - t1):.0f} MB/s", flush=True)
    np.testing.assert_array_equal(v.numpy(), u.numpy())
    assert os.path.getsize(tmp_path / "big_0.cart") == len(_expected_bytes(str(tmp_path), u, mesh, 0.0, np.float64))


def test_iogroup_dump_and_read(ctx, tmp_path):
    """Utils::makeIOGroup over the raw binary streams (IOGroupTest.cpp's dump / read
    pattern with RawBinaryOStream / RawBinaryIStream in place of H5Stream): two fields
    dumped at two time stamps are one file per field and record, each the stream's own
    bytes, and read back in order into fields of the same names."""
    mesh = of.MeshBuilder(ctx).newMesh(17, 9).build()
    u = build_field(ctx, mesh, 2, 1, [(D_, D_), (N_, N_)], name="u")
    v = build_field(ctx, mesh, 2, 1, [(D_, D_), (N_, N_)], name="v")
    u.upload(random_storage(u, SEED + 41))
    v.upload(random_storage(v, SEED + 42))
    u.updatePadding()
    v.updatePadding()
    root = str(tmp_path) + "/"
    g = of.Utils.makeIOGroup(of.Utils.RawBinaryOStream, root, u, v)
    g.dump(0.5)
    u0, v0 = u.numpy(), v.numpy()
    u.assign(u * 2.0 + v)
    v.assign(v - 1.0)
    g.dump(1.0)
    u1, v1 = u.numpy(), v.numpy()
    assert sorted(os.listdir(tmp_path)) == ["u_0.cart", "u_1.cart", "v_0.cart", "v_1.cart"]
    assert cart.read(str(tmp_path / "v_1.cart"))["time"] == 1.0
    with pytest.raises(ValueError):
        g.setAllInOne(True)
    a = build_field(ctx, mesh, 2, 1, [(D_, D_), (N_, N_)], name="u")
    b = build_field(ctx, mesh, 2, 1, [(D_, D_), (N_, N_)], name="v")
    r = of.Utils.makeIOGroup(of.Utils.RawBinaryIStream, root, of.StreamIn, a, b)
    r.read(0.5)
    np.testing.assert_array_equal(a.numpy(), u0)
    np.testing.assert_array_equal(b.numpy(), v0)
    r.read(1.0)
    np.testing.assert_array_equal(a.numpy(), u1)
    np.testing.assert_array_equal(b.numpy(), v1)


def test_iogroup_cpp_frontend(tmp_path):
    """The same round trip through the C++ front-end (tests/cpp/iogroup_example.cpp)."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "opflow_amd", "bin",
                       "iogroup_example")
    assert os.path.exists(exe), "run __graft_entry__.build()"
    r = subprocess.run([exe, str(tmp_path) + "/"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and '"iogroup_roundtrip": true' in r.stdout
    assert sorted(os.listdir(tmp_path)) == ["u_0.cart", "u_1.cart", "v_0.cart", "v_1.cart"]
