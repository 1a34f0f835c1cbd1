"""Uneven batches in one persistent launch (kfr_amd_batch_*): every length of the mixed-length
workload (7-smooth 18..8192 and primes 17..127, the latter through direct prime passes or Bluestein
as the reference's planner decides) in both directions against FP64, graph replays, host pointers,
autotuning, every pass order and LDS padding the autotuner can choose, and f64 (gpu)."""
from __future__ import annotations

import numpy as np
import pytest

import kfr_amd
import oracle
from tests.common import bound_fp64, fp64_dft, uneven_lengths

pytestmark = pytest.mark.gpu


def _bound(n, dtype):
    return bound_fp64(n, dtype)


def _ref(x, inverse, dtype):
    # FP64 numpy for f32 data (fast); the long double oracle for f64 data
    if dtype == np.complex64:
        xx = x.astype(np.complex128)
        return np.fft.ifft(xx) * len(xx) if inverse else np.fft.fft(xx)
    return fp64_dft(x.astype(np.complex128), inverse)


def _check_all(y, x, sizes, inv, dtype):
    off = 0
    worst = (0.0, None)
    for n, d in zip(sizes, inv):
        ref = _ref(x[off:off + n], bool(d), dtype)
        e = oracle.rel_rms(y[off:off + n], ref)
        assert e <= _bound(n, dtype), (n, d, e)
        worst = max(worst, (e, n))
        off += n
    return worst


def _workload(seed, extra, dtype, lengths=None):
    rng = np.random.default_rng(seed)
    L = uneven_lengths() if lengths is None else lengths
    sizes = np.concatenate([L, L, rng.choice(L, extra)]).astype(np.int64)
    rng.shuffle(sizes)
    inv = rng.integers(0, 2, sizes.size).astype(np.int32)
    x = oracle.complex_uniform(seed, int(sizes.sum()), dtype)
    return sizes, inv, x


def test_uneven_lengths_set():
    L = uneven_lengths()
    assert len(L) == 328 and L.min() == 17 and L.max() == 8192


@pytest.mark.parametrize("autotune", [False, True])
def test_batch_every_length_both_directions(autotune):
    torch = pytest.importorskip("torch")
    sizes, inv, x = _workload(3, 600, np.complex64)
    b = kfr_amd.dft_batch(sizes, inv, np.complex64, autotune=autotune)
    assert b.total_size == int(sizes.sum())
    print(b.describe(), flush=True)
    xd = torch.from_numpy(x.view(np.float32).copy()).cuda()
    yd = torch.full_like(xd, float("nan"))
    b.execute(yd, xd)
    torch.cuda.synchronize()
    y = yd.cpu().numpy().view(np.complex64)
    worst = _check_all(y, x, sizes, inv, np.complex64)
    print(f"transforms={sizes.size} worst rel_rms={worst[0]:.3e} at n={worst[1]}", flush=True)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
@pytest.mark.parametrize("order,padshift", [(0, 5), (0, 30), (1, 4), (2, 4), (1, 30), (2, 5)])
def test_batch_forced_pass_orders_and_paddings(order, padshift, dtype, monkeypatch):
    """every schedule the autotuner can choose, forced for all lengths: the planner's pass order (0),
    its reverse (1), odd factors first (2), with LDS padding of one slot per 16 (4), per 32 (5) or none
    (30); every length (f64: up to 1024 points) in both directions against FP64"""
    torch = pytest.importorskip("torch")
    monkeypatch.setenv("KFR_AMD_BATCH_ORDER", str(order))
    monkeypatch.setenv("KFR_AMD_BATCH_PADSHIFT", str(padshift))
    L = uneven_lengths()
    if dtype == np.complex128:
        L = L[L <= 1024]
    sizes, inv, x = _workload(13 + order, 0, dtype, L)
    b = kfr_amd.dft_batch(sizes, inv, dtype)
    info = b.describe()
    print(info, flush=True)
    rdt = np.float32 if dtype == np.complex64 else np.float64
    xd = torch.from_numpy(x.view(rdt).copy()).cuda()
    yd = torch.full_like(xd, float("nan"))
    b.execute(yd, xd)
    torch.cuda.synchronize()
    _check_all(yd.cpu().numpy().view(dtype), x, sizes, inv, dtype)
    pads = [int(v) for v in info.split("padding=")[1].split("/")]
    assert pads[{4: 0, 5: 1, 30: 2}[padshift]] == len(set(sizes.tolist()))


def test_batch_rejects_unknown_padding(monkeypatch):
    monkeypatch.setenv("KFR_AMD_BATCH_PADSHIFT", "3")
    with pytest.raises(kfr_amd.KfrError, match="PADSHIFT"):
        kfr_amd.dft_batch([16, 32], None, np.complex64)


def test_batch_graph_replays_match_first_run():
    """the queue reset and kernel are one captured graph: replays with eager work and
    synchronisations in between give identical results"""
    torch = pytest.importorskip("torch")
    sizes, inv, x = _workload(5, 200, np.complex64)
    b = kfr_amd.dft_batch(sizes, inv, np.complex64)
    xd = torch.from_numpy(x.view(np.float32).copy()).cuda()
    yd = torch.empty_like(xd)
    b.execute(yd, xd)
    torch.cuda.synchronize()
    first = yd.cpu().numpy().copy()
    _check_all(first.view(np.complex64), x, sizes, inv, np.complex64)
    for r in range(4):
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        yd.fill_(float(r + 1))
        torch.cuda.synchronize()
        b.execute(yd, xd)
        torch.cuda.synchronize()
        assert float(junk) == float(junk)
        assert np.array_equal(yd.cpu().numpy(), first), r


def test_batch_new_buffers_and_in_place():
    torch = pytest.importorskip("torch")
    sizes, inv, x = _workload(7, 50, np.complex64)
    b = kfr_amd.dft_batch(sizes, inv, np.complex64)
    xd = torch.from_numpy(x.view(np.float32).copy()).cuda()
    y1 = torch.empty_like(xd)
    b.execute(y1, xd)
    b.execute(xd, xd)  # another pair of pointers: captured again, in place
    torch.cuda.synchronize()
    assert np.array_equal(xd.cpu().numpy(), y1.cpu().numpy())


def test_batch_host_pointers():
    sizes, inv, x = _workload(9, 20, np.complex64)
    b = kfr_amd.dft_batch(sizes, inv, np.complex64)
    y = np.empty_like(x)
    b.execute(y, x)
    _check_all(y, x, sizes, inv, np.complex64)


def test_batch_f64():
    torch = pytest.importorskip("torch")
    L = uneven_lengths()
    L = L[L <= 1024]
    sizes, inv, x = _workload(11, 100, np.complex128, L)
    b = kfr_amd.dft_batch(sizes, inv, np.complex128)
    xd = torch.from_numpy(x.view(np.float64).copy()).cuda()
    yd = torch.empty_like(xd)
    b.execute(yd, xd)
    torch.cuda.synchronize()
    _check_all(yd.cpu().numpy().view(np.complex128), x, sizes, inv, np.complex128)


def test_batch_rejects_oversized_length():
    with pytest.raises(kfr_amd.KfrError, match="does not fit"):
        kfr_amd.dft_batch([16, 1 << 16], None, np.complex64)
