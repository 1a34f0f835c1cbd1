"""Differentiable batched FFT convolution for PyTorch over the C ABI.

`fft_convolve(x, h)` computes y[b] = x[b] * h[b] (linear convolution, len nx + nh - 1) with the GPU
real-FFT convolution, and `fft_correlate(x, h)` the correlation (the convolution with h reversed, as the
reference's correlate); their backward pass calls the hand-written adjoint kernels: the signal gradient
is the correlation of the incoming gradient with the taps and the tap gradient is the correlation
with the signal, both through the conjugate spectrum, with the signal spectrum recomputed rather than
stored. Supports float32 and float64 CUDA tensors, h of shape (nh,) (shared) or (B, nh).
"""
from __future__ import annotations

import torch

from ._lib import check, lib

_plans: dict = {}


def _plan(dtype, nx, nh):
    key = (dtype, nx, nh)
    p = _plans.get(key)
    if p is None:
        p = lib().kfr_amd_conv_create(0 if dtype == torch.float32 else 1, nx, nh)
        check(bool(p))
        _plans[key] = p
    return p


def _stream():
    return torch.cuda.current_stream().cuda_stream


class FFTConvolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x: torch.Tensor, h: torch.Tensor, correlate: bool = False):
        if x.dtype not in (torch.float32, torch.float64) or h.dtype != x.dtype:
            raise TypeError("x and h must both be float32 or both float64")
        if not (x.is_cuda and h.is_cuda):
            raise ValueError("fft_convolve runs on the GPU: move x and h to cuda")
        x2 = x.contiguous().reshape(-1, x.shape[-1])
        shared = h.dim() == 1
        h2 = h.contiguous().reshape(-1, h.shape[-1])
        B, nx = x2.shape
        nh = h2.shape[-1]
        if not shared and h2.shape[0] != B:
            raise ValueError("h must have shape (nh,) or (batch, nh)")
        y = torch.empty(B, nx + nh - 1, dtype=x.dtype, device=x.device)
        p = _plan(x.dtype, nx, nh)
        r = lib().kfr_amd_conv_forward(p, y.data_ptr(), x2.data_ptr(), h2.data_ptr(), 0 if shared else nh, B,
                                       int(bool(correlate)), _stream())
        check(r == 0)
        ctx.save_for_backward(x2, h2)
        ctx.shared = shared
        ctx.correlate = bool(correlate)
        ctx.xshape = x.shape
        ctx.hshape = h.shape
        return y.reshape(x.shape[:-1] + (nx + nh - 1,))

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        x2, h2 = ctx.saved_tensors
        B, nx = x2.shape
        nh = h2.shape[-1]
        g2 = g.contiguous().reshape(B, nx + nh - 1)
        dx = torch.empty_like(x2) if ctx.needs_input_grad[0] else None
        dh = torch.empty(nh if ctx.shared else (B, nh), dtype=x2.dtype, device=x2.device) \
            if ctx.needs_input_grad[1] else None
        p = _plan(x2.dtype, nx, nh)
        r = lib().kfr_amd_conv_backward(p, dx.data_ptr() if dx is not None else None,
                                        dh.data_ptr() if dh is not None else None, g2.data_ptr(), x2.data_ptr(),
                                        h2.data_ptr(), 0 if ctx.shared else nh, B, int(ctx.correlate), _stream())
        check(r == 0)
        return (dx.reshape(ctx.xshape) if dx is not None else None,
                dh.reshape(ctx.hshape) if dh is not None else None, None)


def fft_convolve(x: torch.Tensor, h: torch.Tensor) -> torch.Tensor:
    return FFTConvolve.apply(x, h, False)


def fft_correlate(x: torch.Tensor, h: torch.Tensor) -> torch.Tensor:
    """y[b][m] = sum_j x[b][m - j] h[b][nh - 1 - j]: the correlation of include/kfr/dft/convolution.hpp:62-67
    (convolution with h reversed), differentiable."""
    return FFTConvolve.apply(x, h, True)
