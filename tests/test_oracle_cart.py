"""Pin of the oracle's `.cart` restatement (CPU only): the bytes of a small file,
spelled out item by item in the order the reference's documented structure lists
them (RawBinaryStream.hpp:84-97: name, dim, total ranks, time stamp, mesh range,
mesh coordinates, field global (accessible) range, field local range, data),
and a read of them back."""
import struct

import numpy as np

from oracle import cart


def test_cart_bytes_follow_the_documented_structure(tmp_path):
    path = str(tmp_path / "u_0.cart")
    x0, x1 = np.linspace(0, 1, 4), np.linspace(-1, 1, 3)
    vals = np.arange(6, dtype=np.float64).reshape(2, 3) * 0.5  # local range 3 (x) by 2 (y)
    cart.write(path, "u", 0.25, [(0, 4), (0, 3)], [x0, x1], [(0, 3), (0, 2)], [(0, 3), (0, 2)], vals)
    want = b"".join([
        (1).to_bytes(4, "little"), b"u",                      # std::string name as length + bytes
        (2).to_bytes(4, "little"), (1).to_bytes(4, "little"),  # dim, ranks
        struct.pack("<d", 0.25),                              # time stamp
        struct.pack("<4i", 0, 4, 0, 3),                       # mesh range, start/end per dim
        x0.astype("<f8").tobytes(), x1.astype("<f8").tobytes(),
        struct.pack("<4i", 0, 3, 0, 2), struct.pack("<4i", 0, 3, 0, 2),
        struct.pack("<6d", 0.0, 0.5, 1.0, 1.5, 2.0, 2.5),     # x-fastest
    ])
    assert open(path, "rb").read() == want
    r = cart.read(path)
    assert r["name"] == "u" and r["dim"] == 2 and r["nproc"] == 1 and r["time"] == 0.25
    assert r["mesh_range"] == [(0, 4), (0, 3)] and r["local_range"] == [(0, 3), (0, 2)]
    np.testing.assert_array_equal(r["mesh_x"][1], x1)
    np.testing.assert_array_equal(r["values"], vals.reshape(-1))


def test_cart_float_values(tmp_path):
    path = str(tmp_path / "v_3.cart")
    vals = np.arange(8, dtype=np.float32)
    cart.write(path, "v", 1.0, [(0, 3)] * 3, [np.zeros(3)] * 3, [(0, 2)] * 3, [(0, 2)] * 3, vals, np.float32)
    raw = open(path, "rb").read()
    assert raw[-32:] == vals.astype("<f4").tobytes()
    np.testing.assert_array_equal(cart.read(path, np.float32)["values"], vals)
