"""Test infrastructure: the checkers for the GPU DFT engine.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this package, and only
as the checker (or, for the CPU baseline, as the timed reference). The product never imports it.

* `restated` — the plain-C restatement of the reference path in long double (kfr_oracle.c);
* `reference()` — the reference DFT engine itself, compiled from its own sources by
  oracle/Makefile into oracle/_ref/ (absent when that build was not possible).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_oracle = None
_ref = None
_ref_tried = False

vp = C.c_void_p
sz = C.c_size_t


def _load_oracle():
    global _oracle
    if _oracle is None:
        p = HERE / "libkfr_oracle.so"
        if not p.exists():
            raise ImportError(f"{p} missing: run `make -C oracle`")
        L = C.CDLL(str(p))
        L.oracle_dft.argtypes = [sz, vp, vp, C.c_int]
        L.oracle_rdft.argtypes = [sz, vp, vp, C.c_int]
        L.oracle_irdft.argtypes = [sz, vp, vp, C.c_int]
        L.oracle_dct.argtypes = [sz, vp, vp, C.c_int]
        L.oracle_convolve.argtypes = [vp, sz, vp, sz, vp, C.c_int, C.c_int, C.c_int]
        L.oracle_chirp.argtypes = [sz, vp]
        L.oracle_bitrev.argtypes = [sz, vp]
        L.oracle_digitrev.argtypes = [vp, sz, vp]
        for f in ("oracle_dft", "oracle_rdft", "oracle_irdft", "oracle_dct", "oracle_convolve", "oracle_chirp",
                  "oracle_bitrev", "oracle_digitrev"):
            getattr(L, f).restype = None
        _oracle = L
    return _oracle


def _c128(x):
    return np.ascontiguousarray(x, dtype=np.complex128)


def _f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


class restated:
    """long double restatement (oracle/kfr_oracle.c)."""

    @staticmethod
    def dft(x, inverse=False):
        x = _c128(x)
        out = np.empty_like(x)
        _load_oracle().oracle_dft(x.size, x.ctypes.data, out.ctypes.data, int(inverse))
        return out

    @staticmethod
    def rdft(x, fmt=1):
        x = _f64(x)
        n = x.size
        out = np.empty(n // 2 + 1 if fmt == 1 else n // 2, dtype=np.complex128)
        _load_oracle().oracle_rdft(n, x.ctypes.data, out.ctypes.data, fmt)
        return out

    @staticmethod
    def irdft(X, n, fmt=1):
        X = _c128(X)
        out = np.empty(n, dtype=np.float64)
        _load_oracle().oracle_irdft(n, X.ctypes.data, out.ctypes.data, fmt)
        return out

    @staticmethod
    def dct(x, inverse=False):
        x = _f64(x)
        out = np.empty_like(x)
        _load_oracle().oracle_dct(x.size, x.ctypes.data, out.ctypes.data, int(inverse))
        return out

    @staticmethod
    def convolve(a, b, correlate=False, keep_real_only=True):
        cplx = np.iscomplexobj(a) or np.iscomplexobj(b)
        if cplx:
            a, b = _c128(a), _c128(b)
            out = np.empty(a.size + b.size - 1, dtype=np.complex128)
        else:
            a, b = _f64(a), _f64(b)
            out = np.empty(a.size + b.size - 1, dtype=np.float64)
        _load_oracle().oracle_convolve(a.ctypes.data, a.size, b.ctypes.data, b.size, out.ctypes.data, int(correlate),
                                       int(cplx), int(keep_real_only))
        return out

    @staticmethod
    def chirp(n):
        out = np.empty(n, dtype=np.complex128)
        _load_oracle().oracle_chirp(n, out.ctypes.data)
        return out

    @staticmethod
    def bitrev(n):
        out = np.empty(n, dtype=np.int64)
        _load_oracle().oracle_bitrev(n, out.ctypes.data)
        return out

    @staticmethod
    def digitrev(radices):
        r = np.ascontiguousarray(radices, dtype=np.int32)
        out = np.empty(int(np.prod(radices)), dtype=np.int64)
        _load_oracle().oracle_digitrev(r.ctypes.data, r.size, out.ctypes.data)
        return out


def _cpu_has_avx512() -> bool:
    try:
        flags = Path("/proc/cpuinfo").read_text()
    except OSError:
        return False
    return all(f in flags for f in ("avx512f", "avx512bw", "avx512dq", "avx512vl"))


class _Ref:
    """The reference engine (oracle/_ref/libkfr_ref_<arch>.so) through oracle/kfr_ref_shim.cpp."""

    def __init__(self, path):
        L = C.CDLL(str(path))
        L.kref_version.restype = C.c_char_p
        L.kref_describe.argtypes = [C.c_int, C.c_int, sz, C.c_int, C.c_char_p, sz]
        L.kref_permutation.argtypes = [C.c_int, sz, vp]
        L.kref_dft.argtypes = [C.c_int, sz, sz, vp, vp, C.c_int, C.c_int]
        L.kref_rdft.argtypes = [C.c_int, sz, C.c_int, sz, vp, vp, C.c_int, C.c_int]
        L.kref_dct.argtypes = [C.c_int, sz, sz, vp, vp, C.c_int, C.c_int]
        L.kref_convolve.argtypes = [C.c_int, C.c_int, vp, sz, vp, sz, vp]
        L.kref_convolve.restype = C.c_long
        L.kref_filter_create.argtypes = [C.c_int, vp, sz, sz]
        L.kref_filter_create.restype = vp
        L.kref_filter_apply.argtypes = [vp, vp, vp, sz]
        L.kref_filter_reset.argtypes = [vp]
        L.kref_filter_delete.argtypes = [vp]
        L.kref_dft_md.argtypes = [C.c_int, sz, vp, vp, vp, C.c_int]
        L.kref_dft_md.restype = C.c_long
        L.kref_rdft_md.argtypes = [C.c_int, sz, vp, C.c_int, vp, vp, C.c_int]
        L.kref_rdft_md.restype = C.c_long
        L.kref_time.argtypes = [C.c_int, C.c_int, sz, sz, C.c_int, C.c_int, C.c_int, vp, vp]
        L.kref_time.restype = C.c_double
        self.L = L
        self.path = Path(path)

    def version(self) -> str:
        return self.L.kref_version().decode()

    def describe(self, kind, dtype, n, arg=0) -> str:
        buf = C.create_string_buffer(1 << 16)
        self.L.kref_describe(kind, dtype, n, arg, buf, len(buf))
        return buf.value.decode()

    def permutation(self, dtype, n):
        out = np.empty(n, dtype=np.int64)
        r = self.L.kref_permutation(dtype, n, out.ctypes.data)
        return out if r == 1 else None

    def dft(self, x, inverse=False, threads=1):
        x = np.ascontiguousarray(x)
        out = np.empty_like(x)
        n = x.shape[-1]
        batch = x.size // n
        dt = 0 if x.dtype == np.complex64 else 1
        self.L.kref_dft(dt, n, batch, out.ctypes.data, x.ctypes.data, int(inverse), threads)
        return out

    def rdft(self, x, fmt=1, threads=1):
        x = np.ascontiguousarray(x)
        n = x.shape[-1]
        batch = x.size // n
        cs = n // 2 + 1 if fmt == 1 else n // 2
        cdt = np.complex64 if x.dtype == np.float32 else np.complex128
        out = np.empty(x.shape[:-1] + (cs,), dtype=cdt)
        self.L.kref_rdft(0 if x.dtype == np.float32 else 1, n, fmt, batch, out.ctypes.data, x.ctypes.data, 0, threads)
        return out

    def irdft(self, X, n, fmt=1, threads=1):
        X = np.ascontiguousarray(X)
        cs = X.shape[-1]
        batch = X.size // cs
        rdt = np.float32 if X.dtype == np.complex64 else np.float64
        out = np.empty(X.shape[:-1] + (n,), dtype=rdt)
        self.L.kref_rdft(0 if rdt == np.float32 else 1, n, fmt, batch, out.ctypes.data, X.ctypes.data, 1, threads)
        return out

    def dct(self, x, inverse=False, threads=1):
        x = np.ascontiguousarray(x)
        out = np.empty_like(x)
        n = x.shape[-1]
        self.L.kref_dct(0 if x.dtype == np.float32 else 1, n, x.size // n, out.ctypes.data, x.ctypes.data,
                        int(inverse), threads)
        return out

    def dft_md(self, x, inverse=False):
        """(output, temp_size) of the reference dft_plan_md over every axis of x."""
        x = np.ascontiguousarray(x)
        out = np.empty_like(x)
        sh = np.ascontiguousarray(x.shape, dtype=np.uint32)
        t = self.L.kref_dft_md(0 if x.dtype == np.complex64 else 1, x.ndim, sh.ctypes.data, out.ctypes.data,
                               x.ctypes.data, int(inverse))
        return out, t

    def rdft_md(self, x, real_out_is_enough=False):
        """(complex output, temp_size) of the reference dft_plan_md_real, forward."""
        x = np.ascontiguousarray(x)
        sh = np.ascontiguousarray(x.shape, dtype=np.uint32)
        cdt = np.complex64 if x.dtype == np.float32 else np.complex128
        out = np.empty(x.shape[:-1] + (x.shape[-1] // 2 + 1,), dtype=cdt)
        t = self.L.kref_rdft_md(0 if x.dtype == np.float32 else 1, x.ndim, sh.ctypes.data, int(real_out_is_enough),
                                out.ctypes.data, x.ctypes.data, 0)
        return out, t

    def irdft_md(self, X, shape, real_out_is_enough=False):
        """(real output, temp_size) of the reference dft_plan_md_real, inverse (unscaled)."""
        X = np.ascontiguousarray(X)
        sh = np.ascontiguousarray(shape, dtype=np.uint32)
        rdt = np.float32 if X.dtype == np.complex64 else np.float64
        out = np.empty(X.size * 2, dtype=rdt)  # room for real_out_size()
        t = self.L.kref_rdft_md(0 if rdt == np.float32 else 1, len(shape), sh.ctypes.data, int(real_out_is_enough),
                                out.ctypes.data, X.ctypes.data, 1)
        return out[: int(np.prod(shape))].reshape(shape), t

    def convolve(self, a, b=None, mode=0):
        a = np.ascontiguousarray(a)
        dt = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.complex64): 2,
              np.dtype(np.complex128): 3}[a.dtype]
        b = a if b is None else np.ascontiguousarray(b, dtype=a.dtype)
        n = self.L.kref_convolve(dt, mode, a.ctypes.data, a.size, b.ctypes.data, b.size, None)
        out = np.empty(n, dtype=a.dtype)
        self.L.kref_convolve(dt, mode, a.ctypes.data, a.size, b.ctypes.data, b.size, out.ctypes.data)
        return out

    def filter(self, taps, block=1024):
        taps = np.ascontiguousarray(taps)
        dt = {np.dtype(np.float32): 0, np.dtype(np.float64): 1}[taps.dtype]
        h = self.L.kref_filter_create(dt, taps.ctypes.data, taps.size, block)
        L = self.L

        class F:
            def apply(self_, x):
                x = np.ascontiguousarray(x, dtype=taps.dtype)
                out = np.empty_like(x)
                L.kref_filter_apply(h, out.ctypes.data, x.ctypes.data, x.size)
                return out

            def reset(self_):
                L.kref_filter_reset(h)

            def __del__(self_):
                L.kref_filter_delete(h)

        return F()

    def time(self, kind, dtype, n, batch, inverse, threads, reps, out, inp) -> float:
        return self.L.kref_time(kind, dtype, n, batch, int(inverse), threads, reps, out.ctypes.data, inp.ctypes.data)


def reference(arch: str | None = None):
    """The compiled reference engine for this CPU, or None when oracle/_ref holds no usable build."""
    global _ref, _ref_tried
    if arch is None and _ref_tried:
        return _ref
    cand = []
    if arch:
        cand = [HERE / "_ref" / f"libkfr_ref_{arch}.so"]
    else:
        if _cpu_has_avx512():
            cand.append(HERE / "_ref" / "libkfr_ref_avx512.so")
        cand.append(HERE / "_ref" / "libkfr_ref_avx2.so")
    r = None
    for p in cand:
        if p.exists():
            try:
                r = _Ref(p)
                break
            except OSError:
                continue
    if arch is None:
        _ref, _ref_tried = r, True
    return r


# ------------------------------------------------------------------------------------------------
# Seeded inputs shared by the tests, the smoke check and the benchmark: splitmix64 counters,
# uniform values from the top 24 (f32) or 53 (f64) bits, Gaussian values by Box-Muller in FP64.

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix(seed: int, count: int) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (np.arange(1, count + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)) + np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, count: int, dtype=np.float32) -> np.ndarray:
    """i.i.d. uniform [-1, 1)"""
    z = _splitmix(seed, count)
    if np.dtype(dtype) == np.float32:
        u = (z >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))
        return (2.0 * u - 1.0).astype(np.float32)
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    return 2.0 * u - 1.0


def gaussian(seed: int, count: int, dtype=np.float32) -> np.ndarray:
    """i.i.d. standard normal (Box-Muller in FP64, rounded once)"""
    m = (count + 1) // 2
    z = _splitmix(seed, 2 * m)
    u1 = ((z[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / (1 << 53))
    u2 = (z[1::2] >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    r = np.sqrt(-2.0 * np.log(u1))
    g = np.empty(2 * m, dtype=np.float64)
    g[0::2] = r * np.cos(2 * np.pi * u2)
    g[1::2] = r * np.sin(2 * np.pi * u2)
    return g[:count].astype(dtype)


def complex_uniform(seed: int, shape, dtype=np.complex64) -> np.ndarray:
    n = int(np.prod(shape))
    rdt = np.float32 if np.dtype(dtype) == np.complex64 else np.float64
    v = uniform(seed, 2 * n, rdt)
    return v.view(dtype).reshape(shape)


def complex_gaussian(seed: int, shape, dtype=np.complex64) -> np.ndarray:
    n = int(np.prod(shape))
    rdt = np.float32 if np.dtype(dtype) == np.complex64 else np.float64
    v = gaussian(seed, 2 * n, rdt)
    return v.view(dtype).reshape(shape)


def rel_rms(y, ref) -> float:
    """||y - ref||_2 / ||ref||_2 in FP64"""
    y = np.asarray(y, dtype=np.complex128 if np.iscomplexobj(y) or np.iscomplexobj(ref) else np.float64)
    ref = np.asarray(ref, dtype=y.dtype)
    d = np.linalg.norm((y - ref).ravel())
    r = np.linalg.norm(ref.ravel())
    return float(d / r) if r > 0 else float(d)
