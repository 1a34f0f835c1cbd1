"""`.npy` reading and writing against the reference's own fixture files (tests/npy, copied into
tests/golden/npy): the checks of tests/unit/base/npy.cpp:50-132 (every file loads as the counter
0, 1, 2, ... in C order with its shape; saving that counter gives the version-1 little-endian
C-order file byte for byte), plus the decode errors."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from kfr_amd import npy

DIR = Path(__file__).resolve().parent / "golden" / "npy"
TYPES = ["uint8", "uint16", "uint32", "uint64", "int8", "int16", "int32", "int64", "float32", "float64",
         "complex64", "complex128"]
SHAPES = [(110,), (10, 11), (2, 5, 11)]


def _name(shape, t, ver, order, endian):
    return "x".join(str(s) for s in shape) + f"-{t}-v{ver}-{order}-{endian}e.npy"


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("t", TYPES)
def test_reference_files(shape, t):
    ref = np.arange(int(np.prod(shape))).astype(t).reshape(shape)
    checked = 0
    for ver in (1, 2):
        for order in ("c", "f"):
            for endian in ("l", "b"):
                f = DIR / _name(shape, t, ver, order, endian)
                if not f.exists():
                    continue  # one-byte types have a single byte order
                code, arr = npy.load_bytes(f.read_bytes(), t, len(shape))
                assert code == npy.npy_decode_result.ok, (f.name, code)
                assert arr.shape == shape and arr.dtype == np.dtype(t)
                assert np.array_equal(arr, ref), f.name
                checked += 1
                if ver == 1 and order == "c" and endian == "l":
                    assert npy.save_bytes(ref) == f.read_bytes(), f.name
    assert checked >= 4


def test_decode_errors():
    good = npy.save_bytes(np.arange(6, dtype=np.float32).reshape(2, 3))
    assert npy.load_bytes(good, np.float32, 2)[0] == npy.npy_decode_result.ok
    assert npy.load_bytes(good, np.float64, 2)[0] == npy.npy_decode_result.invalid_type
    assert npy.load_bytes(good, np.float32, 1)[0] == npy.npy_decode_result.invalid_shape
    assert npy.load_bytes(b"\x93NUMPX" + good[6:], np.float32, 2)[0] == npy.npy_decode_result.cannot_read
    assert npy.load_bytes(good[:-4], np.float32, 2)[0] == npy.npy_decode_result.cannot_read
    bad = good.replace(b"'shape'", b"'shape!").replace(b"(2, 3)", b"[2, 3]")
    assert npy.load_bytes(bad, np.float32, 2)[0] == npy.npy_decode_result.invalid_header
    empty = npy.save_bytes(np.zeros((0, 3), dtype=np.float32))
    assert npy.load_bytes(empty, np.float32, 2)[0] == npy.npy_decode_result.invalid_shape


def test_round_trip_and_numpy_agreement():
    x = (np.arange(24) * (1 + 0.5j)).astype(np.complex64).reshape(2, 3, 4)
    b = npy.save_bytes(x)
    assert len(b) % 64 == (x.nbytes % 64)  # data starts at a multiple of 64
    code, y = npy.load_bytes(b, np.complex64, 3)
    assert code == npy.npy_decode_result.ok and np.array_equal(x, y)
    import io
    assert np.array_equal(np.load(io.BytesIO(b)), x)
    assert npy.save_bytes(np.asfortranarray(x)) == b  # saved in C order whatever the layout
