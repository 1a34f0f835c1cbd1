"""ctypes binding of libkfr_amd.so (declarations in include/kfr_amd_capi.h).

The library is the product: when it is missing this module raises instead of falling back to any
CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libkfr_amd.so"
HEADER = PKG.parent / "include" / "kfr_amd_capi.h"

_lib = None

c_size = C.c_size_t
vp = C.c_void_p


class KfrError(RuntimeError):
    """An entry point failed; the message is kfr_last_error()."""


def _proto(lib):
    def f(name, res, *args):
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = list(args)

    f("kfr_version_string", C.c_char_p)
    f("kfr_version", C.c_uint32)
    f("kfr_enabled_archs", C.c_char_p)
    f("kfr_current_arch", C.c_int)
    f("kfr_last_error", C.c_char_p)
    f("kfr_allocate", vp, c_size)
    f("kfr_allocate_aligned", vp, c_size, c_size)
    f("kfr_deallocate", None, vp)
    for t in ("f32", "f64"):
        f(f"kfr_dft_create_plan_{t}", vp, c_size)
        f(f"kfr_dft_create_2d_plan_{t}", vp, c_size, c_size)
        f(f"kfr_dft_create_3d_plan_{t}", vp, c_size, c_size, c_size)
        f(f"kfr_dft_create_md_plan_{t}", vp, c_size, vp)
        f(f"kfr_dft_dump_{t}", None, vp)
        f(f"kfr_dft_get_size_{t}", c_size, vp)
        f(f"kfr_dft_get_temp_size_{t}", c_size, vp)
        f(f"kfr_dft_execute_{t}", None, vp, vp, vp, vp)
        f(f"kfr_dft_execute_inverse_{t}", None, vp, vp, vp, vp)
        f(f"kfr_dft_delete_plan_{t}", None, vp)
        f(f"kfr_dft_real_create_plan_{t}", vp, c_size, C.c_int)
        f(f"kfr_dft_real_create_2d_plan_{t}", vp, c_size, c_size, C.c_bool)
        f(f"kfr_dft_real_create_3d_plan_{t}", vp, c_size, c_size, c_size, C.c_bool)
        f(f"kfr_dft_real_create_md_plan_{t}", vp, c_size, vp, C.c_bool)
        f(f"kfr_dft_real_dump_{t}", None, vp)
        f(f"kfr_dft_real_get_size_{t}", c_size, vp)
        f(f"kfr_dft_real_get_temp_size_{t}", c_size, vp)
        f(f"kfr_dft_real_execute_{t}", None, vp, vp, vp, vp)
        f(f"kfr_dft_real_execute_inverse_{t}", None, vp, vp, vp, vp)
        f(f"kfr_dft_real_delete_plan_{t}", None, vp)
        f(f"kfr_dct_create_plan_{t}", vp, c_size)
        f(f"kfr_dct_dump_{t}", None, vp)
        f(f"kfr_dct_get_size_{t}", c_size, vp)
        f(f"kfr_dct_get_temp_size_{t}", c_size, vp)
        f(f"kfr_dct_execute_{t}", None, vp, vp, vp, vp)
        f(f"kfr_dct_execute_inverse_{t}", None, vp, vp, vp, vp)
        f(f"kfr_dct_delete_plan_{t}", None, vp)
        f(f"kfr_filter_create_fir_plan_{t}", vp, vp, c_size)
        f(f"kfr_filter_create_iir_plan_{t}", vp, vp, c_size)
        f(f"kfr_filter_create_convolution_plan_{t}", vp, vp, c_size, c_size)
        f(f"kfr_filter_process_{t}", None, vp, vp, vp, c_size)
        f(f"kfr_filter_reset_{t}", None, vp)
        f(f"kfr_filter_delete_plan_{t}", None, vp)
        f(f"kfr_amd_dft_execute_batch_{t}", C.c_int, vp, vp, vp, c_size, C.c_int, vp)
        f(f"kfr_amd_dft_real_execute_batch_{t}", C.c_int, vp, vp, vp, c_size, C.c_int, vp)
        f(f"kfr_amd_dct_execute_batch_{t}", C.c_int, vp, vp, vp, c_size, C.c_int, vp)
        f(f"kfr_amd_dft_real_get_complex_size_{t}", c_size, vp)
    for t in ("c32", "c64"):
        f(f"kfr_amd_filter_create_convolution_plan_{t}", vp, vp, c_size, c_size)
        f(f"kfr_amd_filter_process_{t}", None, vp, vp, vp, c_size)
        f(f"kfr_amd_filter_reset_{t}", None, vp)
        f(f"kfr_amd_filter_delete_plan_{t}", None, vp)
    for t in ("f32", "f64", "c32", "c64"):
        f(f"kfr_amd_filter_set_data_{t}", None, vp, vp, c_size)
    for t in ("f32", "f64"):
        f(f"kfr_amd_cache_dft_plan_{t}", vp, c_size)
        f(f"kfr_amd_cache_dft_real_plan_{t}", vp, c_size, C.c_int)
        f(f"kfr_amd_dft_progressive_start_{t}", vp, vp, C.c_int, C.c_int, vp, vp)
        f(f"kfr_amd_dft_real_progressive_start_{t}", vp, vp, C.c_int, vp, vp)
    f("kfr_amd_cache_clear", None)
    f("kfr_amd_cache_size", c_size)
    f("kfr_amd_progressive_total_steps", c_size, vp)
    f("kfr_amd_progressive_step", C.c_int, vp)
    f("kfr_amd_progressive_delete", None, vp)
    f("kfr_amd_dft_progressive_steps", c_size, C.c_int, c_size, C.c_int)
    f("kfr_amd_dft_real_progressive_steps", c_size, C.c_int, c_size, C.c_int)
    f("kfr_amd_dft_execute_batch_noalloc_f32", C.c_int, vp, vp, vp, c_size, C.c_int, vp)
    f("kfr_amd_convolve", C.c_long, C.c_int, C.c_int, vp, c_size, vp, c_size, vp)
    f("kfr_amd_conv_create", vp, C.c_int, c_size, c_size)
    f("kfr_amd_conv_fft_size", c_size, vp)
    f("kfr_amd_conv_forward", C.c_int, vp, vp, vp, vp, c_size, c_size, C.c_int, vp)
    f("kfr_amd_conv_backward", C.c_int, vp, vp, vp, vp, vp, vp, c_size, c_size, C.c_int, vp)
    f("kfr_amd_conv_delete", None, vp)
    f("kfr_amd_describe", C.c_int, C.c_int, C.c_int, c_size, C.c_int, C.c_char_p, c_size)
    f("kfr_amd_dft_permutation", C.c_int, C.c_int, c_size, vp)
    f("kfr_amd_dft_progressive_total_steps", c_size, C.c_int, c_size)
    f("kfr_amd_dft_schedule", C.c_int, C.c_int, c_size, C.c_char_p, c_size)
    f("kfr_amd_batch_create", vp, C.c_int, c_size, vp, vp, C.c_int)
    f("kfr_amd_batch_total_size", c_size, vp)
    f("kfr_amd_batch_execute", C.c_int, vp, vp, vp, vp)
    f("kfr_amd_batch_describe", C.c_int, vp, C.c_char_p, c_size)
    f("kfr_amd_batch_delete", None, vp)
    f("kfr_amd_dft_half_create", vp, C.c_int, c_size)
    f("kfr_amd_dft_half_execute_batch", C.c_int, vp, vp, vp, c_size, C.c_int, vp)
    f("kfr_amd_dft_half_describe", C.c_int, vp, C.c_char_p, c_size)
    f("kfr_amd_dft_half_delete", None, vp)
    f("kfr_amd_inject_alloc_failure", None, C.c_int)
    f("kfr_amd_synchronize", C.c_int)
    f("kfr_amd_timing_begin", C.c_int)
    f("kfr_amd_timing_end", C.c_longlong)
    f("kfr_amd_timing_get", C.c_int, C.c_longlong, C.c_char_p, c_size, vp, vp)


def lib():
    """The loaded library. Raises if it has not been built: there is no fallback."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m kfr_amd.build` "
                              "(the DFT engine has no CPU fallback)")
        l = C.CDLL(str(LIB_PATH))
        _proto(l)
        _lib = l
    return _lib


def last_error() -> str:
    return lib().kfr_last_error().decode()


def check(ok: bool = True):
    """Raise KfrError if the last call set an error."""
    e = last_error()
    if e or not ok:
        raise KfrError(e or "kfr_amd call failed")


def declared_symbols() -> list[str]:
    """Every function name include/kfr_amd_capi.h declares."""
    text = HEADER.read_text()
    return sorted(set(re.findall(r"KFR_AMD_API\s+[\w\s\*]+?\b(kfr_\w+)\s*\(", text)))
