"""Half-precision storage (kfr_amd_dft_half_*): bfloat16 and binary16 complex data, FP32 arithmetic,
radix-16 passes on MFMA. Errors are measured against the FP64 transform of the stored (already
rounded) input, so they cover the computation and the rounding of the output (gpu)."""
from __future__ import annotations

import numpy as np
import pytest

import kfr_amd

pytestmark = pytest.mark.gpu

# rel-RMS bounds: the rounding of the output alone is ~1.5e-3 (bfloat16) and ~2e-4 (binary16)
BOUND = {"bf16": 2e-3, "f16": 5e-4}


def _run(storage, n, batch, inverse, scale=1.0, seed=0, in_place=False):
    torch = pytest.importorskip("torch")
    dt = torch.bfloat16 if storage == "bf16" else torch.float16
    g = torch.Generator(device="cuda")
    g.manual_seed(seed + n)
    x = (torch.randn(batch, n, 2, device="cuda", generator=g) * scale).to(dt)
    p = kfr_amd.dft_plan_half(n, storage)
    y = x if in_place else torch.empty_like(x)
    xin = x.clone() if in_place else x
    p.execute_batch(y, x, batch, inverse)
    torch.cuda.synchronize()
    xc = torch.view_as_complex(xin.double().contiguous()).cpu().numpy()
    yc = torch.view_as_complex(y.double().contiguous()).cpu().numpy()
    ref = np.fft.ifft(xc, axis=-1) * n if inverse else np.fft.fft(xc, axis=-1)
    rel = np.sqrt(np.sum(np.abs(yc - ref) ** 2, axis=-1) / np.sum(np.abs(ref) ** 2, axis=-1))
    agg = float(np.sqrt(np.sum(np.abs(yc - ref) ** 2) / np.sum(np.abs(ref) ** 2)))
    mx = float(np.max(np.abs(yc - ref)) / np.max(np.abs(ref)))
    return float(rel.max()), mx, p.describe(), agg


@pytest.mark.parametrize("storage", ["bf16", "f16"])
@pytest.mark.parametrize("L", [4, 5, 7, 8, 10, 12, 14, 15, 17, 20])
def test_half_sizes(storage, L):
    n = 1 << L
    batch = max(1, min(64, (1 << 21) >> L)) + (1 if L < 14 else 0)
    for inverse in (False, True):
        rel, mx, d, agg = _run(storage, n, batch, inverse, seed=L)
        print(f"{storage} 2^{L} batch={batch} inv={inverse} rel_rms batch={agg:.3e} worst transform={rel:.3e} "
              f"max={mx:.3e} {d}", flush=True)
        # the output rounding of a short transform scatters around its mean: the batch meets the bound,
        # single transforms of a few points may exceed it a little
        assert agg <= BOUND[storage], (storage, L, inverse, agg)
        assert rel <= 1.6 * BOUND[storage], (storage, L, inverse, rel)


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_half_config_2p20_batch64(storage):
    rel, mx, d, _ = _run(storage, 1 << 20, 64, False, seed=99)
    print(f"{storage} 2^20 x64 rel_rms={rel:.3e} max={mx:.3e} {d}", flush=True)
    assert rel <= BOUND[storage]


def test_half_intermediate_scaling_keeps_precision_across_ranges():
    """bfloat16 data far outside binary16's range: the per-group scale of the binary16 intermediate
    keeps the relative error unchanged"""
    for scale in (1e-30, 1e30):
        rel, mx, _, _ = _run("bf16", 1 << 16, 4, False, scale=scale, seed=5)
        print(f"scale={scale:g} rel_rms={rel:.3e}", flush=True)
        assert rel <= BOUND["bf16"]


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_half_in_place(storage):
    for L in (10, 18):
        rel, _, _, _ = _run(storage, 1 << L, 3, False, seed=L, in_place=True)
        assert rel <= BOUND[storage]


def test_half_rejects_bad_arguments():
    with pytest.raises(kfr_amd.KfrError, match="power-of-two"):
        kfr_amd.dft_plan_half(1000, "bf16")
    L = kfr_amd.lib()
    assert not L.kfr_amd_dft_half_create(3, 1024)
    assert "storage" in kfr_amd.last_error()
