"""Kernel timing records (kfr_amd_timing_*): names, algorithmic bytes and durations of the launches
made inside a timing block (gpu), and the empty block (cpu)."""
from __future__ import annotations

import numpy as np
import pytest

import kfr_amd


def test_timing_empty_block():
    with kfr_amd.kernel_timing() as kt:
        pass
    assert kt.records == [] and not kt.by_kernel()
    L = kfr_amd.lib()
    assert L.kfr_amd_timing_get(0, None, 0, None, None) == -1
    assert "out of range" in kfr_amd.last_error()


@pytest.mark.gpu
def test_timing_records_launches():
    torch = pytest.importorskip("torch")
    n, batch = 4096, 96
    x = torch.randn(batch * n * 2, device="cuda")
    y = torch.empty_like(x)
    p = kfr_amd.dft_plan(n, np.complex64)
    p.execute_batch(y, x, batch, False)
    torch.cuda.synchronize()
    with kfr_amd.kernel_timing() as kt:
        for _ in range(3):
            p.execute_batch(y, x, batch, False)
        p.execute_batch(y, x, batch, True)
    k = kt.by_kernel()
    print(k, flush=True)
    fwd, inv = "k_pow2<float, 12, false, 0, 0>", "k_pow2<float, 12, true, 0, 0>"
    assert set(k) == {fwd, inv}
    assert k[fwd]["launches"] == 3 and k[inv]["launches"] == 1
    assert k[fwd]["bytes"] == 3 * 2 * 8 * n * batch
    assert all(ms > 0 for _, _, ms in kt.records)
    # records stop with the block
    p.execute_batch(y, x, batch, False)
    torch.cuda.synchronize()
    assert len(kt.records) == 4


@pytest.mark.gpu
def test_timing_four_step_kernels_cover_the_transform():
    """a four-step length records its two kernels once per round, and the bytes add up to two
    reads and two writes of the data"""
    torch = pytest.importorskip("torch")
    n, batch = 1 << 20, 4
    x = torch.randn(batch * n * 2, device="cuda")
    y = torch.empty_like(x)
    p = kfr_amd.dft_plan(n, np.complex64)
    with kfr_amd.kernel_timing() as kt:
        p.execute_batch(y, x, batch, False)
    names = [r[0] for r in kt.records]
    print(names, flush=True)
    assert len(names) >= 2 and all(nm.startswith("k_pow2<float, ") for nm in names)
    assert sum(r[1] for r in kt.records) == 2 * 2 * 8 * n * batch
