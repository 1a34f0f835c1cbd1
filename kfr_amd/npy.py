"""`.npy` reading and writing with the reference's conventions (include/kfr/base/npy.hpp:36-470), so
that fixtures and results round-trip with KFR's own tooling.

    save_to_npy(array, write)          -> None    write(bytes) receives the file in pieces
    load_from_npy(read, dtype, ndim)   -> (npy_decode_result, array or None)
                                                  read(n) returns n bytes (fewer: cannot_read)

Writing follows save_to_npy (npy.hpp:361-392): format version 1.0, the header
"{'descr': '<f4', 'fortran_order': False, 'shape': (10, 11), }" padded with spaces and a newline so
that the data starts at a multiple of 64 bytes, then the values in C order. Reading follows
load_from_npy (npy.hpp:417-468): versions 1 to 3, the element type must be the requested one in
either byte order ('<' or '>', converted), the number of dimensions must match, empty shapes are
rejected, and Fortran-ordered data is returned with Fortran strides.
"""
from __future__ import annotations

import enum
import struct

import numpy as np

MAGIC = b"\x93NUMPY"


class npy_decode_result(enum.IntEnum):
    """npy.hpp:397-404"""
    ok = 0
    cannot_read = 1
    invalid_header = 2
    invalid_type = 3
    invalid_shape = 4


def npy_format(dtype) -> str:
    """The type string written for an element type (npy.hpp:38-73): one-byte types carry '|'."""
    dt = np.dtype(dtype)
    if dt == np.bool_:
        return "|b1"
    kind = {"f": "f", "u": "u", "i": "i", "c": "c"}.get(dt.kind)
    if kind is None:
        raise TypeError(f"no .npy type for {dt}")
    return ("|" if dt.itemsize == 1 else "<") + kind + str(dt.itemsize)


def _shape_text(shape) -> str:
    """npy.hpp:75-87: '(110,)' for one dimension, '(10, 11)' otherwise."""
    body = ", ".join(str(int(s)) for s in shape)
    return "(" + body + ("," if len(shape) == 1 else "") + ")"


def save_to_npy(array, write) -> None:
    a = np.ascontiguousarray(array)
    write(MAGIC)
    write(struct.pack("<H", 0x0001))
    header = "{'descr': '" + npy_format(a.dtype) + "', 'fortran_order': False, 'shape': " + _shape_text(a.shape) + ", }"
    written = len(MAGIC) + 2
    total = -(-(written + 2 + len(header) + 1) // 64) * 64
    header += " " * (total - (written + 2 + len(header) + 1)) + "\n"
    write(struct.pack("<H", len(header)))
    write(header.encode("latin1"))
    write(a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes())


def save_bytes(array) -> bytes:
    parts: list[bytes] = []
    save_to_npy(array, parts.append)
    return b"".join(parts)


# ---- header dictionary parsing (npy.hpp:153-349)

class _Cursor:
    def __init__(self, text: str):
        self.s = text
        self.i = 0

    def skip_ws(self) -> bool:
        while self.i < len(self.s) and ord(self.s[self.i]) <= ord(" "):
            self.i += 1
        return self.i < len(self.s)

    def peek(self) -> str:
        return self.s[self.i] if self.i < len(self.s) else ""

    def string(self, q: str = "'"):
        if self.peek() != q:
            return None
        self.i += 1
        out = []
        while self.i < len(self.s) and self.s[self.i] != q:
            if self.s[self.i] == "\\":
                self.i += 1
                if self.i >= len(self.s):
                    return None
            out.append(self.s[self.i])
            self.i += 1
        if self.i >= len(self.s):
            return None
        self.i += 1
        return "".join(out)

    def boolean(self):
        if self.s.startswith("False", self.i):
            self.i += 5
            return False
        if self.s.startswith("True", self.i):
            self.i += 4
            return True
        return None

    def integer(self):
        j = self.i
        while self.i < len(self.s) and self.s[self.i].isdigit():
            self.i += 1
        return int(self.s[j:self.i]) if self.i > j else None

    def shape(self):
        if self.peek() != "(":
            return None
        self.i += 1
        if not self.skip_ws():
            return None
        dims = []
        while True:
            n = self.integer()
            if n is None:
                return None
            dims.append(n)
            if not self.skip_ws():
                return None
            if self.peek() == ")":
                self.i += 1
                return dims
            if self.peek() != ",":
                return None
            self.i += 1
            if not self.skip_ws():
                return None
            if self.peek() == ")":
                self.i += 1
                return dims

    def skip_value(self) -> bool:
        ctx = ""
        while self.i < len(self.s):
            c = self.s[self.i]
            if c in "'\"":
                if self.string(c) is None:
                    return False
                continue
            if c in "([{":
                ctx += c
            elif c in ")]}":
                if not ctx:
                    return c == "}"
                ctx = ctx[:-1]
            elif c == "," and not ctx:
                return True
            self.i += 1
        return False


def _decode_dict(text: str):
    c = _Cursor(text)
    if len(text) < 10 or c.peek() != "{":
        return None
    c.i += 1
    hdr = {"descr": None, "fortran_order": False, "shape": None}
    while c.i < len(c.s) and c.peek() != "}":
        if not c.skip_ws():
            return None
        name = c.string()
        if name is None or not c.skip_ws() or c.peek() != ":":
            return None
        c.i += 1
        if not c.skip_ws():
            return None
        if name == "descr":
            v = c.string()
        elif name == "fortran_order":
            v = c.boolean()
        elif name == "shape":
            v = c.shape()
        else:
            v = True if c.skip_value() else None
        if v is None:
            return None
        if name in hdr:
            hdr[name] = v
        if c.peek() == "}":
            return hdr
        if c.peek() != ",":
            return None
        c.i += 1
        if not c.skip_ws():
            return None
    return hdr if c.peek() == "}" else None


def load_from_npy(read, dtype, ndim: int):
    def take(n: int):
        b = read(n)
        return b if b is not None and len(b) == n else None

    if take(len(MAGIC)) != MAGIC:
        return npy_decode_result.cannot_read, None
    v = take(2)
    if v is None:
        return npy_decode_result.cannot_read, None
    (ver,) = struct.unpack("<H", v)
    if ver not in (1, 2, 3):
        return npy_decode_result.cannot_read, None
    hl = take(4 if ver >= 2 else 2)
    if hl is None:
        return npy_decode_result.cannot_read, None
    (header_len,) = struct.unpack("<I" if ver >= 2 else "<H", hl)
    header = take(header_len)
    if header is None:
        return npy_decode_result.cannot_read, None
    hdr = _decode_dict(header.decode("latin1"))
    if hdr is None or hdr["descr"] is None or hdr["shape"] is None:
        return npy_decode_result.invalid_header, None
    fmt = npy_format(dtype)
    descr = hdr["descr"]
    swap = False
    if descr != fmt:
        if len(descr) > 1 and descr[0] != fmt[0] and descr[1:] == fmt[1:]:
            swap = descr[0] == ">"
        else:
            return npy_decode_result.invalid_type, None
    shape = hdr["shape"]
    if len(shape) != ndim:
        return npy_decode_result.invalid_shape, None
    count = int(np.prod(shape))
    if count == 0:
        return npy_decode_result.invalid_shape, None
    dt = np.dtype(dtype)
    data = take(count * dt.itemsize)
    if data is None:
        return npy_decode_result.cannot_read, None
    flat = np.frombuffer(data, dtype=dt.newbyteorder(">" if swap else "<")).astype(dt)
    arr = flat.reshape(shape, order="F" if hdr["fortran_order"] else "C")
    return npy_decode_result.ok, arr


def load_bytes(blob: bytes, dtype, ndim: int):
    pos = [0]

    def read(n):
        b = blob[pos[0]:pos[0] + n]
        pos[0] += n
        return b

    return load_from_npy(read, dtype, ndim)
