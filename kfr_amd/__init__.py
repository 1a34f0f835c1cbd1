"""MI355X (gfx950) implementation of KFR's DFT engine.

The engine is libkfr_amd.so (HIP kernels behind a C ABI that matches the reference's kfr_dft_* /
kfr_dft_real_* / kfr_dct_* / kfr_filter_* functions). This package is a thin Python binding over
that ABI: the class API of the reference's C++ headers, and a torch.autograd.Function for FFT
convolution with hand-written adjoint kernels (kfr_amd.autograd).
"""
from ._lib import KfrError, last_error, lib
from .dft import (CCs, Perm, autocorrelate, convolve, convolve_filter, correlate, dct_plan, describe, dft,
                  dft_batch, dft_cache, dft_plan_half, dft_plan_md, dft_plan_md_real, dft_plan, dft_plan_real, idft, irealdft,
                  permutation, realdft, schedule)
from .timing import kernel_timing

__all__ = ["KfrError", "last_error", "lib", "CCs", "Perm", "autocorrelate", "convolve", "convolve_filter",
           "correlate", "dct_plan", "describe", "dft", "dft_batch", "dft_cache", "dft_plan_half", "dft_plan_md", "dft_plan_md_real", "dft_plan", "dft_plan_real", "idft", "irealdft", "realdft",
           "kernel_timing", "permutation", "schedule"]
