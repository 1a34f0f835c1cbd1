"""Python mirror of the reference's DFT class API over the C ABI.

Names, argument meaning and error behaviour follow include/kfr/dft/fft.hpp and
include/kfr/dft/convolution.hpp: dft_plan, dft_plan_real, dct_plan, convolve, correlate,
autocorrelate, convolve_filter and dft_cache. Arrays may be numpy arrays (host memory, copied to the
GPU and back) or torch tensors on the GPU (used in place). No scaling is applied by any transform.
"""
from __future__ import annotations

import ctypes as C
import sys
import threading

import numpy as np

from ._lib import KfrError, check, lib

CCs = 1
Perm = 0


def _ptr(a):
    """Data pointer of a numpy array or a torch tensor."""
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return a.data_ptr()
    raise TypeError(f"unsupported array type {type(a)!r}")


def _nbytes(a):
    if isinstance(a, np.ndarray):
        return a.nbytes
    return a.numel() * a.element_size()


def _tag(dtype) -> str:
    dt = np.dtype(dtype)
    if dt in (np.float32, np.complex64):
        return "f32"
    if dt in (np.float64, np.complex128):
        return "f64"
    raise TypeError("only float32/float64 (and their complex types) are supported")


def _stream_ptr(stream):
    """A hipStream_t handle: the given torch stream or handle, else torch's current stream when torch
    is in use, else the default stream (NULL)."""
    if stream is None:
        torch = sys.modules.get("torch")
        if torch is not None and torch.cuda.is_available():
            return torch.cuda.current_stream().cuda_stream or None
        return None
    h = getattr(stream, "cuda_stream", stream)
    return h or None


class progressive:
    """State of a progressive execution (dft_plan::progressive, include/kfr/dft/fft.hpp:364-374): a native
    handle that keeps the plan and the buffers alive until the last step."""

    def __init__(self, handle, keep):
        self._h = handle
        self._keep = keep
        self.step = 0
        self.total = int(lib().kfr_amd_progressive_total_steps(handle))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().kfr_amd_progressive_delete(h)
            self._h = None


def _progressive_step(prog: progressive) -> bool:
    r = lib().kfr_amd_progressive_step(prog._h)
    check(r >= 0)
    prog.step += 1
    return r == 1


class dft_plan:
    """Complex DFT of `size` points (include/kfr/dft/fft.hpp:183-456)."""

    def __init__(self, size: int, dtype=np.complex64, progressive_optimized: bool = False, _handle=None):
        self.t = _tag(dtype)
        self._progressive_optimized = progressive_optimized
        L = lib()
        self._h = _handle if _handle is not None else getattr(L, f"kfr_dft_create_plan_{self.t}")(size)
        if not self._h:
            raise KfrError(L.kfr_last_error().decode())
        self.size = getattr(L, f"kfr_dft_get_size_{self.t}")(self._h)
        self.temp_size = getattr(L, f"kfr_dft_get_temp_size_{self.t}")(self._h)

    @property
    def complex_dtype(self):
        return np.complex64 if self.t == "f32" else np.complex128

    def execute(self, out, in_, temp=None, inverse: bool = False):
        fn = f"kfr_dft_execute_inverse_{self.t}" if inverse else f"kfr_dft_execute_{self.t}"
        getattr(lib(), fn)(self._h, _ptr(out), _ptr(in_), None if temp is None else _ptr(temp))
        check()
        return out

    def execute_batch(self, out, in_, batch: int, inverse: bool = False, stream=None):
        """Batched transform of device tensors (contiguous items of `size` points), asynchronous."""
        r = getattr(lib(), f"kfr_amd_dft_execute_batch_{self.t}")(self._h, _ptr(out), _ptr(in_), batch,
                                                                   int(inverse), _stream_ptr(stream))
        check(r == 0)
        return out

    def dump(self):
        getattr(lib(), f"kfr_dft_dump_{self.t}")(self._h)

    def describe(self) -> str:
        return describe(3 if self._progressive_optimized else 0, self.t, self.size)

    def progressive_total_steps(self) -> int:
        return lib().kfr_amd_dft_progressive_steps(0 if self.t == "f32" else 1, self.size,
                                                    int(self._progressive_optimized))

    # progressive execution (include/kfr/dft/fft.hpp:393-439): each step runs its share of the
    # transform's GPU launches synchronously; `out` holds the result after the last step
    def progressive_start(self, inverse, out, in_, temp=None) -> progressive:
        if self.temp_size and temp is None:
            raise KfrError("Temporary buffer must be provided for progressive execution")
        h = getattr(lib(), f"kfr_amd_dft_progressive_start_{self.t}")(self._h, int(self._progressive_optimized),
                                                                      int(bool(inverse)), _ptr(out), _ptr(in_))
        check(bool(h))
        return progressive(h, (self, out, in_))

    def progressive_step(self, prog: progressive) -> bool:
        return _progressive_step(prog)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            getattr(lib(), f"kfr_dft_delete_plan_{self.t}")(h)
            self._h = None


class dft_plan_real:
    """Real DFT of `size` (even) points (include/kfr/dft/fft.hpp:459-566)."""

    def __init__(self, size: int, fmt: int = CCs, dtype=np.float32, _handle=None):
        self.t = _tag(dtype)
        L = lib()
        self._h = _handle if _handle is not None else getattr(L, f"kfr_dft_real_create_plan_{self.t}")(size, fmt)
        if not self._h:
            raise KfrError(L.kfr_last_error().decode())
        self.size = getattr(L, f"kfr_dft_real_get_size_{self.t}")(self._h)
        self.temp_size = getattr(L, f"kfr_dft_real_get_temp_size_{self.t}")(self._h)
        self.fmt = fmt

    def complex_size(self) -> int:
        return self.size // 2 + 1 if self.fmt == CCs else self.size // 2

    def execute(self, out, in_, temp=None):
        """Forward when `in_` is real (writes complex_size() values), inverse when it is complex."""
        is_complex = np.dtype(in_.dtype).kind == "c" if isinstance(in_, np.ndarray) else in_.is_complex()
        fn = f"kfr_dft_real_execute_inverse_{self.t}" if is_complex else f"kfr_dft_real_execute_{self.t}"
        getattr(lib(), fn)(self._h, _ptr(out), _ptr(in_), None if temp is None else _ptr(temp))
        check()
        return out

    def execute_batch(self, out, in_, batch: int, inverse: bool = False, stream=None):
        r = getattr(lib(), f"kfr_amd_dft_real_execute_batch_{self.t}")(self._h, _ptr(out), _ptr(in_), batch,
                                                                        int(inverse), _stream_ptr(stream))
        check(r == 0)
        return out

    def progressive_total_steps(self) -> int:
        return lib().kfr_amd_dft_real_progressive_steps(0 if self.t == "f32" else 1, self.size, self.fmt)

    def progressive_start(self, out, in_, temp=None) -> progressive:
        """Forward when `in_` is real, inverse when it is complex (include/kfr/dft/fft.hpp:538-556)."""
        if self.temp_size and temp is None:
            raise KfrError("Temporary buffer must be provided for progressive execution")
        is_complex = np.dtype(in_.dtype).kind == "c" if isinstance(in_, np.ndarray) else in_.is_complex()
        h = getattr(lib(), f"kfr_amd_dft_real_progressive_start_{self.t}")(self._h, int(is_complex), _ptr(out),
                                                                           _ptr(in_))
        check(bool(h))
        return progressive(h, (self, out, in_))

    def progressive_step(self, prog: progressive) -> bool:
        return _progressive_step(prog)

    def dump(self):
        getattr(lib(), f"kfr_dft_real_dump_{self.t}")(self._h)

    def describe(self) -> str:
        return describe(1, self.t, self.size, self.fmt)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            getattr(lib(), f"kfr_dft_real_delete_plan_{self.t}")(h)
            self._h = None


class dft_plan_md:
    """Complex DFT over every axis of a row-major array (include/kfr/dft/fft.hpp:569-691), through
    kfr_dft_create_md_plan_*; executed with the same kfr_dft_execute_* calls as one-axis plans."""

    def __init__(self, shape, dtype=np.complex64):
        self.t = _tag(dtype)
        self.shape = tuple(int(v) for v in shape)
        sh = (C.c_uint * len(self.shape))(*self.shape)
        L = lib()
        self._h = getattr(L, f"kfr_dft_create_md_plan_{self.t}")(len(self.shape), C.cast(sh, C.c_void_p))
        if not self._h:
            raise KfrError(L.kfr_last_error().decode())
        self.size = getattr(L, f"kfr_dft_get_size_{self.t}")(self._h)
        self.temp_size = getattr(L, f"kfr_dft_get_temp_size_{self.t}")(self._h)

    def execute(self, out, in_, temp=None, inverse: bool = False):
        fn = f"kfr_dft_execute_inverse_{self.t}" if inverse else f"kfr_dft_execute_{self.t}"
        getattr(lib(), fn)(self._h, _ptr(out), _ptr(in_), None if temp is None else _ptr(temp))
        check()
        return out

    def execute_batch(self, out, in_, batch: int, inverse: bool = False, stream=None):
        """`batch` arrays of device tensors, one after another, on `stream` (asynchronous)."""
        r = getattr(lib(), f"kfr_amd_dft_execute_batch_{self.t}")(self._h, _ptr(out), _ptr(in_), batch,
                                                                   int(inverse), _stream_ptr(stream))
        check(r == 0)
        return out

    def dump(self):
        getattr(lib(), f"kfr_dft_dump_{self.t}")(self._h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            getattr(lib(), f"kfr_dft_delete_plan_{self.t}")(h)
            self._h = None


class dft_plan_md_real:
    """Real DFT over every axis of a row-major array (include/kfr/dft/fft.hpp:693-924): real input of
    `shape`, complex output of `complex_size()` (the last axis n/2+1, CCs), and back."""

    def __init__(self, shape, real_out_is_enough: bool = False, dtype=np.float32):
        self.t = _tag(dtype)
        self.shape = tuple(int(v) for v in shape)
        self.real_out_is_enough = bool(real_out_is_enough)
        sh = (C.c_uint * len(self.shape))(*self.shape)
        L = lib()
        self._h = getattr(L, f"kfr_dft_real_create_md_plan_{self.t}")(len(self.shape), C.cast(sh, C.c_void_p),
                                                                      self.real_out_is_enough)
        if not self._h:
            raise KfrError(L.kfr_last_error().decode())
        self.size = getattr(L, f"kfr_dft_real_get_size_{self.t}")(self._h)
        self.temp_size = getattr(L, f"kfr_dft_real_get_temp_size_{self.t}")(self._h)

    def complex_size(self) -> tuple:
        return self.shape[:-1] + (self.shape[-1] // 2 + 1,)

    def execute(self, out, in_, temp=None):
        """Forward when `in_` is real, inverse (real output) when it is complex."""
        is_complex = np.dtype(in_.dtype).kind == "c" if isinstance(in_, np.ndarray) else in_.is_complex()
        fn = f"kfr_dft_real_execute_inverse_{self.t}" if is_complex else f"kfr_dft_real_execute_{self.t}"
        getattr(lib(), fn)(self._h, _ptr(out), _ptr(in_), None if temp is None else _ptr(temp))
        check()
        return out

    def dump(self):
        getattr(lib(), f"kfr_dft_real_dump_{self.t}")(self._h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            getattr(lib(), f"kfr_dft_real_delete_plan_{self.t}")(h)
            self._h = None


class dct_plan:
    """DCT-II (inverse: DCT-III with x[0] halved), unscaled (include/kfr/dft/fft.hpp:927-978)."""

    def __init__(self, size: int, dtype=np.float32):
        self.t = _tag(dtype)
        L = lib()
        self._h = getattr(L, f"kfr_dct_create_plan_{self.t}")(size)
        if not self._h:
            raise KfrError(L.kfr_last_error().decode())
        self.size = getattr(L, f"kfr_dct_get_size_{self.t}")(self._h)
        self.temp_size = getattr(L, f"kfr_dct_get_temp_size_{self.t}")(self._h)

    def execute(self, out, in_, temp=None, inverse: bool = False):
        fn = f"kfr_dct_execute_inverse_{self.t}" if inverse else f"kfr_dct_execute_{self.t}"
        getattr(lib(), fn)(self._h, _ptr(out), _ptr(in_), None if temp is None else _ptr(temp))
        check()
        return out

    def execute_batch(self, out, in_, batch: int, inverse: bool = False, stream=None):
        r = getattr(lib(), f"kfr_amd_dct_execute_batch_{self.t}")(self._h, _ptr(out), _ptr(in_), batch,
                                                                  int(inverse), _stream_ptr(stream))
        check(r == 0)
        return out

    def dump(self):
        getattr(lib(), f"kfr_dct_dump_{self.t}")(self._h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            getattr(lib(), f"kfr_dct_delete_plan_{self.t}")(h)
            self._h = None


def _conv(mode: int, a: np.ndarray, b: np.ndarray | None):
    a = np.ascontiguousarray(a)
    dt = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.complex64): 2,
          np.dtype(np.complex128): 3}[a.dtype]
    if b is None:
        b = a
    b = np.ascontiguousarray(b, dtype=a.dtype)
    L = lib()
    n = L.kfr_amd_convolve(dt, mode, a.ctypes.data, a.size, b.ctypes.data, b.size, None)
    check(n >= 0)
    out = np.empty(n, dtype=a.dtype)
    r = L.kfr_amd_convolve(dt, mode, a.ctypes.data, a.size, b.ctypes.data, b.size, out.ctypes.data)
    check(r >= 0)
    return out


def convolve(a, b):
    """Linear convolution, len(a)+len(b)-1 values (include/kfr/dft/convolution.hpp:53-59)."""
    return _conv(0, a, b)


def correlate(a, b):
    """Correlation: convolution with b reversed (include/kfr/dft/convolution.hpp:62-67)."""
    return _conv(1, a, b)


def autocorrelate(a):
    """Second half of correlate(a, a) (include/kfr/dft/convolution.hpp:70-76)."""
    return _conv(2, a, None)


class convolve_filter:
    """Streaming FFT convolution filter (include/kfr/dft/convolution.hpp:99-148), real or complex taps
    (complex through the kfr_amd_filter_*_c32/c64 extensions)."""

    def __init__(self, taps, block_size: int = 1024):
        taps = np.ascontiguousarray(taps)
        self.t = _tag(taps.dtype)
        if taps.dtype.kind == "c":
            self.t = "c32" if self.t == "f32" else "c64"
        self.dtype = taps.dtype
        L = lib()
        self._create = (f"kfr_amd_filter_create_convolution_plan_{self.t}" if self.t[0] == "c"
                        else f"kfr_filter_create_convolution_plan_{self.t}")
        self._process = f"kfr_amd_filter_process_{self.t}" if self.t[0] == "c" else f"kfr_filter_process_{self.t}"
        self._reset = f"kfr_amd_filter_reset_{self.t}" if self.t[0] == "c" else f"kfr_filter_reset_{self.t}"
        self._delete = (f"kfr_amd_filter_delete_plan_{self.t}" if self.t[0] == "c"
                        else f"kfr_filter_delete_plan_{self.t}")
        self._h = getattr(L, self._create)(taps.ctypes.data, taps.size, block_size)
        if not self._h:
            raise KfrError(L.kfr_last_error().decode())
        b = 1
        while b < (block_size or 1024):
            b <<= 1
        self.block_size = b

    def input_block_size(self) -> int:
        return self.block_size

    def apply(self, dest, src=None):
        if src is None:
            src = dest
        n = min(dest.size if isinstance(dest, np.ndarray) else dest.numel(),
                src.size if isinstance(src, np.ndarray) else src.numel())
        getattr(lib(), self._process)(self._h, _ptr(dest), _ptr(src), n)
        check()
        return dest

    def reset(self):
        getattr(lib(), self._reset)(self._h)
        check()

    def set_data(self, taps):
        """Replace the filter taps and reset the state (src/dft/convolution-impl.cpp:58-84)."""
        taps = np.ascontiguousarray(taps, dtype=self.dtype)
        getattr(lib(), f"kfr_amd_filter_set_data_{self.t}")(self._h, taps.ctypes.data, taps.size)
        check()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            getattr(lib(), self._delete)(h)
            self._h = None


class dft_batch:
    """Many complex transforms of uneven lengths and per-transform directions in one launch
    (kfr_amd_batch_*). Transform i reads and writes sizes[i] values starting at the sum of the sizes
    before it. No counterpart in the reference, which plans one length at a time."""

    def __init__(self, sizes, inverse=None, dtype=np.complex64, autotune: bool = False):
        self.dtype = np.dtype(dtype)
        self.sizes = np.ascontiguousarray(sizes, dtype=np.uint64)
        inv = None if inverse is None else np.ascontiguousarray(inverse, dtype=np.int32)
        if inv is not None and inv.size != self.sizes.size:
            raise ValueError("inverse must have one flag per transform")
        self._inv = inv
        self._p = lib().kfr_amd_batch_create(0 if _tag(dtype) == "f32" else 1, self.sizes.size,
                                            self.sizes.ctypes.data, None if inv is None else inv.ctypes.data,
                                            1 if autotune else 0)
        check(bool(self._p))

    @property
    def total_size(self) -> int:
        return int(lib().kfr_amd_batch_total_size(self._p))

    def offsets(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.sizes.astype(np.int64))[:-1]])

    def execute(self, out, in_, stream=None):
        """Device tensors are enqueued on the stream (torch's current stream by default); numpy arrays
        are staged synchronously."""
        if _nbytes(out) < self.total_size * self.dtype.itemsize or _nbytes(in_) < self.total_size * self.dtype.itemsize:
            raise ValueError("buffers smaller than the batch")
        s = _stream_ptr(stream) if not isinstance(out, np.ndarray) else None
        check(lib().kfr_amd_batch_execute(self._p, _ptr(out), _ptr(in_), s) == 0)
        return out

    def describe(self) -> str:
        buf = C.create_string_buffer(4096)
        check(lib().kfr_amd_batch_describe(self._p, buf, len(buf)) >= 0)
        return buf.value.decode()

    def __del__(self):
        p = getattr(self, "_p", None)
        if p:
            lib().kfr_amd_batch_delete(p)
            self._p = None


class dft_plan_half:
    """Complex DFT of a power-of-two size on half-precision data (kfr_amd_dft_half_*): pairs of
    bfloat16 (storage "bf16") or binary16 ("f16"), computed in FP32 with the radix-16 passes on MFMA.
    Buffers are torch tensors (torch.bfloat16 / torch.float16, interleaved re, im) or any array of
    4-byte complex values. No counterpart in the reference."""

    def __init__(self, size: int, storage: str = "bf16"):
        self.storage = storage
        code = {"bf16": 1, "f16": 2}[storage]
        self._h = lib().kfr_amd_dft_half_create(code, size)
        check(bool(self._h))
        self.size = int(size)

    def execute_batch(self, out, in_, batch: int, inverse: bool = False, stream=None):
        s = _stream_ptr(stream) if not isinstance(out, np.ndarray) else None
        check(lib().kfr_amd_dft_half_execute_batch(self._h, _ptr(out), _ptr(in_), batch, int(inverse), s) == 0)
        return out

    def describe(self) -> str:
        buf = C.create_string_buffer(1024)
        check(lib().kfr_amd_dft_half_describe(self._h, buf, len(buf)) >= 0)
        return buf.value.decode()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().kfr_amd_dft_half_delete(h)
            self._h = None


class dft_cache:
    """The process-wide plan cache (include/kfr/dft/cache.hpp:44-126) held by the library
    (kfr_amd_cache_*): get() and getreal() return plans that share the cached GPU plan; its key names
    the element type, transform kind, size, pack format, batch layout and schedule variant."""

    _inst = None
    _lock = threading.Lock()

    @classmethod
    def instance(cls) -> "dft_cache":
        with cls._lock:
            if cls._inst is None:
                cls._inst = dft_cache()
            return cls._inst

    def get(self, dtype, size: int) -> dft_plan:
        t = _tag(dtype)
        h = getattr(lib(), f"kfr_amd_cache_dft_plan_{t}")(int(size))
        check(bool(h))
        return dft_plan(size, np.complex64 if t == "f32" else np.complex128, _handle=h)

    def getreal(self, dtype, size: int, fmt: int = CCs) -> dft_plan_real:
        t = _tag(dtype)
        h = getattr(lib(), f"kfr_amd_cache_dft_real_plan_{t}")(int(size), fmt)
        check(bool(h))
        return dft_plan_real(size, fmt, np.float32 if t == "f32" else np.float64, _handle=h)

    def size(self) -> int:
        return int(lib().kfr_amd_cache_size())

    def clear(self):
        lib().kfr_amd_cache_clear()
        check()


def dft(x: np.ndarray) -> np.ndarray:
    """Forward DFT through the cache (include/kfr/dft/cache.hpp:128-137)."""
    x = np.ascontiguousarray(x)
    p = dft_cache.instance().get(x.dtype, x.size)
    out = np.full_like(x, np.nan)
    return p.execute(out, x)


def idft(x: np.ndarray) -> np.ndarray:
    """Inverse DFT through the cache, unscaled (include/kfr/dft/cache.hpp:139-148)."""
    x = np.ascontiguousarray(x)
    p = dft_cache.instance().get(x.dtype, x.size)
    out = np.full_like(x, np.nan)
    return p.execute(out, x, inverse=True)


def realdft(x: np.ndarray) -> np.ndarray:
    """Real forward DFT through the cache, n/2+1 values in CCs format (include/kfr/dft/cache.hpp:150-159)."""
    x = np.ascontiguousarray(x)
    p = dft_cache.instance().getreal(x.dtype, x.size)
    out = np.full(x.size // 2 + 1, np.nan, dtype=np.complex64 if x.dtype == np.float32 else np.complex128)
    return p.execute(out, x)


def irealdft(X: np.ndarray) -> np.ndarray:
    """Real inverse DFT through the cache of a CCs spectrum of (n-1)*2 points, unscaled
    (include/kfr/dft/cache.hpp:161-170)."""
    X = np.ascontiguousarray(X)
    n = (X.size - 1) * 2
    rdt = np.float32 if X.dtype == np.complex64 else np.float64
    p = dft_cache.instance().getreal(rdt, n)
    out = np.full(n, np.nan, dtype=rdt)
    return p.execute(out, X)


def describe(kind: int, t: str, size: int, arg: int = 0) -> str:
    """The reference plan descriptor (dump() layout), computed on the host."""
    buf = C.create_string_buffer(1 << 16)
    n = lib().kfr_amd_describe(kind, 0 if t in ("f32", 0) else 1, size, arg, buf, len(buf))
    check(n >= 0)
    if n >= len(buf):
        buf = C.create_string_buffer(n + 1)
        lib().kfr_amd_describe(kind, 0 if t in ("f32", 0) else 1, size, arg, buf, len(buf))
    return buf.value.decode()


def permutation(t: str, size: int) -> np.ndarray | None:
    """Index array of the reference plan's reorder stage (out[i] = in[perm[i]]), or None."""
    perm = np.empty(size, dtype=np.int64)
    r = lib().kfr_amd_dft_permutation(0 if t == "f32" else 1, size, perm.ctypes.data)
    check(r >= 0)
    return perm if r == 1 else None


def schedule(t: str, size: int) -> str:
    buf = C.create_string_buffer(4096)
    r = lib().kfr_amd_dft_schedule(0 if t == "f32" else 1, size, buf, len(buf))
    check(r >= 0)
    return buf.value.decode()
