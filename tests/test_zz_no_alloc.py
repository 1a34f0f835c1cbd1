"""The allocation-free execute (kfr_amd_dft_execute_batch_noalloc_f32) through the C ABI against the
ordinary batched execute, bit for bit: two calls in a row, the second with a smaller input, on the
four-step path (intermediates in the static workspace, in rounds) and a one-workgroup length; sizes
whose intermediates exceed the workspace are refused with an error (gpu)."""
from __future__ import annotations

import numpy as np
import pytest

import kfr_amd
import oracle

pytestmark = pytest.mark.gpu


def _run(fn, plan, out, x, batch, inverse):
    torch = pytest.importorskip("torch")
    r = fn(plan._h, out.data_ptr(), x.data_ptr(), batch, int(inverse), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return r


@pytest.mark.parametrize("n,b1,b2", [(1 << 20, 12, 3), (1 << 22, 2, 1), (1000, 300, 7), (96 * 1024, 5, 2),
                                     (1024000, 4, 1)])
def test_noalloc_matches_execute_batch(n, b1, b2):
    torch = pytest.importorskip("torch")
    L = kfr_amd.lib()
    plan = kfr_amd.dft_plan(n, np.complex64)
    for batch, seed in ((b1, 1), (b2, 2)):   # the second call is smaller
        x = oracle.complex_uniform(seed + n % 97, (batch, n), np.complex64)
        xd = torch.from_numpy(x.view(np.float32).copy()).cuda()
        for inverse in (False, True):
            y_ref = torch.full_like(xd, float("nan"))
            y_na = torch.full_like(xd, float("nan"))
            assert _run(L.kfr_amd_dft_execute_batch_f32, plan, y_ref, xd, batch, inverse) == 0
            assert _run(L.kfr_amd_dft_execute_batch_noalloc_f32, plan, y_na, xd, batch, inverse) == 0, \
                kfr_amd.last_error()
            assert torch.equal(y_ref, y_na), (n, batch, inverse)
        got = y_na.cpu().numpy().view(np.complex64).reshape(batch, n)[-1]
        assert oracle.rel_rms(got, np.fft.ifft(x[-1].astype(np.complex128)) * n) <= 1e-6


def test_noalloc_refuses_what_does_not_fit():
    torch = pytest.importorskip("torch")
    L = kfr_amd.lib()
    for n in (1 << 23, 1 << 24):
        plan = kfr_amd.dft_plan(n, np.complex64)
        x = torch.zeros(2 * n, device="cuda")
        y = torch.empty_like(x)
        assert _run(L.kfr_amd_dft_execute_batch_noalloc_f32, plan, y, x, 1, False) == -1
        assert "cannot hold" in kfr_amd.last_error()
