"""Mixed-radix kernels specialised per length (f32): every menu length as a single kernel, and
four-step products of two menu lengths (including 15,360,000, the complex half of the 30,720,000-point
real transform), batched on device against FP64 (gpu)."""
from __future__ import annotations

import numpy as np
import pytest

import kfr_amd
import oracle
from tests.common import bound_fp64, fp64_dft

pytestmark = pytest.mark.gpu

MENU = [1000, 1200, 1250, 1280, 1500, 1600, 1792, 1875, 1920, 2000, 2400, 2500, 2560, 3000, 3200, 3584, 3750,
        3840, 4000]
FOURSTEP = [1000 * 1024, 1200 * 1250, 1875 * 4096, 1792 * 3584, 3840 * 4000]


def _check(n, batch, seed):
    torch = pytest.importorskip("torch")
    x = oracle.complex_uniform(seed, (batch, n), np.complex64)
    xd = torch.from_numpy(x.view(np.float32).copy()).cuda()
    yd = torch.empty_like(xd)
    p = kfr_amd.dft_plan(n, np.complex64)
    sched = kfr_amd.schedule("f32", n)
    assert "mixed" in sched, sched
    bound = bound_fp64(n, np.complex64)
    for inverse in (False, True):
        p.execute_batch(yd, xd, batch, inverse)
        torch.cuda.synchronize()
        y = yd.cpu().numpy().view(np.complex64).reshape(batch, n)
        ref = fp64_dft(x.astype(np.complex128), inverse) if n <= 4096 else \
            (np.fft.ifft(x.astype(np.complex128), axis=-1) * n if inverse else np.fft.fft(x.astype(np.complex128), axis=-1))
        errs = [oracle.rel_rms(y[b], ref[b]) for b in range(batch)]
        print(f"n={n} batch={batch} inv={inverse} worst={max(errs):.3e} {sched}", flush=True)
        assert max(errs) <= bound, (n, inverse, max(errs))
    p.execute_batch(xd, xd, batch, False)  # in place
    torch.cuda.synchronize()
    y = xd.cpu().numpy().view(np.complex64).reshape(batch, n)
    assert max(oracle.rel_rms(y[b], np.fft.fft(x[b].astype(np.complex128))) for b in range(batch)) <= bound


@pytest.mark.parametrize("n", MENU)
def test_mixed_single_kernel(n):
    _check(n, 37, n)


@pytest.mark.parametrize("n", FOURSTEP)
def test_mixed_four_step(n):
    _check(n, 2, n % 1000)


def test_real_30720000():
    """dft_plan_real<float>(30,720,000), the reference's own mixed-radix real configuration"""
    torch = pytest.importorskip("torch")
    n = 30720000
    x = np.random.default_rng(11).standard_normal((2, n)).astype(np.float32)
    p = kfr_amd.dft_plan_real(n, kfr_amd.CCs, np.float32)
    xd = torch.from_numpy(x.copy()).cuda()
    yd = torch.empty(2 * (n // 2 + 1) * 2, dtype=torch.float32, device="cuda")
    p.execute_batch(yd, xd, 2, False)
    torch.cuda.synchronize()
    Y = yd.cpu().numpy().view(np.complex64).reshape(2, n // 2 + 1)
    for b in range(2):
        ref = np.fft.rfft(x[b].astype(np.float64))
        e = oracle.rel_rms(Y[b], ref)
        print(f"r2c 30720000 rel_rms={e:.3e}", flush=True)
        assert e <= bound_fp64(n, np.float32)
