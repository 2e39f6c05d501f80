"""Matrix files of VCP (vcp/vcp_portable_fio.hpp, vcp/vcp_fio.hpp) through
the C ABI: round trips of every element type in both formats, the CRC-64
against the CRC-64/ECMA-182 check value, rejection of corrupt files (as the
reference's loader rejects them), and the reference's own data files
(test_PDE/Data_Gray_Scott, kept under tests/golden/vcp_data): read, and
written back byte for byte."""
import os
import struct

import numpy as np
import pytest

import vcp_amd as V

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vcp_data")


def _crc(b):
    import ctypes

    out = ctypes.c_uint64()
    buf = ctypes.create_string_buffer(bytes(b), len(b))
    assert V.lib().vcphip_crc64(buf, len(b), ctypes.byref(out)) == 0
    return out.value


def test_crc64_check_value():
    assert _crc(b"123456789") == 0x6C40DF5F0B497347   # CRC-64/ECMA-182 (MSB first, init 0, no xorout)
    assert _crc(b"") == 0


def _sample(rng, mtype, r, c):
    X = V._mat_array(mtype, r, c)
    if mtype == V.MAT_INT32:
        X[...] = rng.integers(-2 ** 31, 2 ** 31, (r, c))
    else:
        X[...] = rng.standard_normal(X.shape) * 2.0 ** rng.integers(-600, 600, X.shape)
        if X.size:
            X[(0,) * X.ndim] = np.nan
            X[(r - 1, c - 1) + (0,) * (X.ndim - 2)] = -np.inf
    return X


@pytest.mark.parametrize("mtype", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("portable", [True, False])
@pytest.mark.parametrize("r,c", [(1, 1), (7, 3), (1, 9), (40, 1)])
def test_round_trip(tmp_path, mtype, portable, r, c):
    rng = np.random.default_rng(mtype * 100 + r * 10 + c)
    X = _sample(rng, mtype, r, c)
    path = str(tmp_path / "m.bin")
    V.mat_write(path, X, mtype, portable)
    assert V.mat_info(path) == (mtype, r, c, portable)
    Y, t = V.mat_read(path)
    assert t == mtype
    assert np.ascontiguousarray(X).tobytes() == np.ascontiguousarray(Y).tobytes()   # bit for bit, NaN included
    assert not os.path.exists(path + ".tmp")


def test_portable_layout_and_crc(tmp_path):
    X = np.asfortranarray(np.arange(6, dtype=np.float64).reshape(2, 3))
    path = str(tmp_path / "a.vcpmat")
    V.mat_write(path, X, V.MAT_DOUBLE, True)
    b = open(path, "rb").read()
    assert b[:8] == b"VCPMAT2\0"
    assert struct.unpack("<IIII QQQ", b[8:48]) == (2, 2, 1, 0, 2, 3, 48)
    assert struct.unpack("<6d", b[48:96]) == (0.0, 3.0, 1.0, 4.0, 2.0, 5.0)   # column-major
    assert struct.unpack("<Q", b[96:104])[0] == _crc(b[:96]) and len(b) == 104


def test_zero_sized_matrix(tmp_path):
    path = str(tmp_path / "z.vcpmat")
    V.mat_write(path, np.zeros((0, 0)), V.MAT_DOUBLE, True)
    assert V.mat_info(path) == (V.MAT_DOUBLE, 0, 0, True)


@pytest.mark.parametrize("damage", ["flip", "truncate", "trailing", "version", "layout", "type", "size"])
def test_corrupt_portable_files_are_rejected(tmp_path, damage):
    X = np.asfortranarray(np.random.default_rng(1).standard_normal((5, 4)))
    path = str(tmp_path / "c.vcpmat")
    V.mat_write(path, X, V.MAT_DOUBLE, True)
    b = bytearray(open(path, "rb").read())
    if damage == "flip":
        b[60] ^= 0x10
    elif damage == "truncate":
        b = b[:-3]
    elif damage == "trailing":
        b += b"\0"
    elif damage == "version":
        b[8] = 3
    elif damage == "layout":
        b[16] = 2
    elif damage == "type":
        b[12] = 9
    elif damage == "size":
        b[24] = 6
    open(path, "wb").write(bytes(b))
    with pytest.raises(V.VcphipError) as e:
        V.mat_info(path)
    assert e.value.code == -9


def test_type_and_size_mismatch_on_read(tmp_path):
    path = str(tmp_path / "t.vcpmat")
    V.mat_write(path, np.zeros((3, 3)), V.MAT_DOUBLE, True)
    buf = np.zeros((3, 3, 2))
    L = V.lib()
    assert L.vcphip_mat_read(path.encode(), V.MAT_DD, 3, 3, buf.ctypes.data_as(V._vp), 3) == -9
    assert L.vcphip_mat_read(path.encode(), V.MAT_DOUBLE, 3, 4, buf.ctypes.data_as(V._vp), 3) == -1
    assert L.vcphip_mat_info(b"/nonexistent/file", None, None, None, None) == -2


@pytest.mark.parametrize("name", sorted(os.listdir(DATA)))
def test_reference_data_files_read_and_rewrite_byte_for_byte(tmp_path, name):
    path = os.path.join(DATA, name)
    mtype, r, c, portable = V.mat_info(path)
    assert not portable
    want = {"matrix_int": V.MAT_INT32, "matrix_d": V.MAT_DOUBLE, "matrix_kvdd": V.MAT_DD}[name.rsplit(".", 1)[1]]
    assert mtype == want and r > 0 and c > 0
    X, _ = V.mat_read(path)
    if mtype != V.MAT_INT32:
        assert np.isfinite(X).all()
    out = str(tmp_path / name)
    V.mat_write(out, X, mtype, portable=False)
    assert open(out, "rb").read() == open(path, "rb").read()
    # and through the portable format
    V.mat_write(out + ".vcpmat", X, mtype, portable=True)
    Y, t = V.mat_read(out + ".vcpmat")
    assert t == mtype and np.array_equal(X, Y)


@pytest.mark.parametrize("name", sorted(n for n in os.listdir(DATA) if n.endswith(".matrix_kvdd")))
def test_reference_dd_data_is_normalised(name):
    """every kv::dd of the reference's data has |a2| <= ulp(a1)/2"""
    Q, t = V.mat_read(os.path.join(DATA, name))
    assert t == V.MAT_DD
    hi, lo = Q[..., 0], Q[..., 1]
    assert (np.abs(lo) <= np.abs(np.spacing(hi)) * 0.5).all()
    assert (hi + lo == hi).all()
