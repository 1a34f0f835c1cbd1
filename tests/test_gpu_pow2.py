"""Power-of-two C2C kernels (one workgroup per transform group, the four-step split, and the
multi-pass Stockham over HBM for the longest lengths) over contiguous device batches: every length
2^1..2^24, batch tails, in-place, both directions (gpu)."""
from __future__ import annotations

import numpy as np
import pytest

import kfr_amd
import oracle
from tests.common import bound_fp64, fp64_dft

pytestmark = pytest.mark.gpu


def _batch_for(L):
    # odd batch sizes leave a partial last workgroup; about 2^20 values per case
    return max(1, (1 << 20) >> L) + 1 if L <= 18 else (2 if L <= 22 else 1)


@pytest.mark.parametrize("L", list(range(1, 25)))
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_pow2_batch(L, dtype):
    torch = pytest.importorskip("torch")
    n = 1 << L
    batch = _batch_for(L)
    x = oracle.complex_uniform(900 + L, (batch, n), dtype)
    rdt = np.float32 if dtype == np.complex64 else np.float64
    xd = torch.from_numpy(x.view(rdt).copy()).cuda()
    yd = torch.empty_like(xd)
    plan = kfr_amd.dft_plan(n, dtype)
    sched = kfr_amd.schedule("f32" if dtype == np.complex64 else "f64", n)
    bound = bound_fp64(n, dtype)
    ref = {inv: fp64_dft(x.astype(np.complex128), inv) for inv in (False, True)}
    for inverse in (False, True):
        plan.execute_batch(yd, xd, batch, inverse)
        torch.cuda.synchronize()
        y = yd.cpu().numpy().view(dtype).reshape(batch, n)
        errs = [oracle.rel_rms(y[b], ref[inverse][b]) for b in range(batch)]
        print(f"L={L} {np.dtype(dtype).name} batch={batch} inv={inverse} worst={max(errs):.3e} {sched}", flush=True)
        assert max(errs) <= bound, (L, inverse, max(errs), int(np.argmax(errs)))
    # in place, forward
    plan.execute_batch(xd, xd, batch, False)
    torch.cuda.synchronize()
    y = xd.cpu().numpy().view(dtype).reshape(batch, n)
    assert max(oracle.rel_rms(y[b], ref[False][b]) for b in range(batch)) <= bound


def test_pow2_untouched_tail():
    """a batch that ends inside a workgroup writes nothing past its last transform"""
    torch = pytest.importorskip("torch")
    n, batch = 64, 5                       # 64-point f32 transforms run 64 to a workgroup
    x = oracle.complex_uniform(17, (batch, n), np.complex64)
    xd = torch.from_numpy(x.view(np.float32).copy()).cuda()
    guard = torch.full((2 * n * (batch + 3),), 7.0, device="cuda")
    plan = kfr_amd.dft_plan(n, np.complex64)
    plan.execute_batch(guard, xd, batch, False)
    torch.cuda.synchronize()
    g = guard.cpu().numpy()
    assert np.all(g[2 * n * batch:] == 7.0)
    y = g[:2 * n * batch].view(np.complex64).reshape(batch, n)
    assert oracle.rel_rms(y, fp64_dft(x.astype(np.complex128))) <= 1e-6


def test_pow2_f64_four_step_round_trip():
    torch = pytest.importorskip("torch")
    n = 1 << 24
    x = oracle.complex_uniform(23, n, np.complex128)
    plan = kfr_amd.dft_plan(n, np.complex128)
    xd = torch.from_numpy(x.view(np.float64).copy()).cuda()
    y = torch.empty_like(xd)
    plan.execute_batch(y, xd, 1, False)
    plan.execute_batch(y, y, 1, True)
    torch.cuda.synchronize()
    z = y.cpu().numpy().view(np.complex128) / n
    err = oracle.rel_rms(z, x)
    print(f"f64 2^24 round trip rel_rms={err:.3e}", flush=True)
    assert err < 1e-14


def _check_plan(dtype, L, batch, seed):
    torch = pytest.importorskip("torch")
    n = 1 << L
    x = oracle.complex_uniform(seed, (batch, n), dtype)
    rdt = np.float32 if dtype == np.complex64 else np.float64
    xd = torch.from_numpy(x.view(rdt).copy()).cuda()
    yd = torch.empty_like(xd)
    plan = kfr_amd.dft_plan(n, dtype)
    bound = bound_fp64(n, dtype)
    worst = 0.0
    for inverse in (False, True):
        ref = fp64_dft(x.astype(np.complex128), inverse)
        plan.execute_batch(yd, xd, batch, inverse)
        torch.cuda.synchronize()
        y = yd.cpu().numpy().view(dtype).reshape(batch, n)
        worst = max(worst, max(oracle.rel_rms(y[b], ref[b]) for b in range(batch)))
    return worst, bound


# (dtype, log2 n): every four-step split that KFR_AMD_FOURSTEP_L2 can select, including the ones the
# measured table does not pick (rows of 2^14 points, the longest columns)
SPLIT_CASES = [(np.complex64, L) for L in (16, 18, 21, 22)] + [(np.complex128, L) for L in (15, 17, 20, 22)]


@pytest.mark.parametrize("dtype,L", SPLIT_CASES, ids=lambda v: getattr(v, "__name__", str(v)))
def test_pow2_every_fourstep_split(dtype, L, monkeypatch):
    t = "f32" if dtype == np.complex64 else "f64"
    ml = 14 if dtype == np.complex64 else 13
    batch = 3 if L <= 18 else 1
    for l2 in range(max(7, L - ml), min(ml, L - 7) + 1):
        monkeypatch.setenv("KFR_AMD_FOURSTEP_L2", str(l2))
        sched = kfr_amd.schedule(t, 1 << L)
        assert f"n2={1 << l2} " in sched, sched
        worst, bound = _check_plan(dtype, L, batch, 1300 + L + l2)
        print(f"L={L} {t} L2={l2} worst={worst:.3e} {sched}", flush=True)
        assert worst <= bound, (L, l2, worst)


@pytest.mark.parametrize("dtype,L", [(np.complex64, L) for L in (16, 17, 18, 20)] +
                         [(np.complex128, L) for L in (15, 16, 17, 19)],
                         ids=lambda v: getattr(v, "__name__", str(v)))
def test_pow2_multipass_below_threshold(dtype, L, monkeypatch):
    """the multi-pass Stockham selected through KFR_AMD_MULTIPASS_MIN_LOG for lengths above one
    workgroup (passes of radix 2^5..2^8: 2^17 runs 2^5 x 2^6 x 2^6)"""
    t = "f32" if dtype == np.complex64 else "f64"
    monkeypatch.setenv("KFR_AMD_MULTIPASS_MIN_LOG", str(L))
    sched = kfr_amd.schedule(t, 1 << L)
    assert "multi-pass" in sched, sched
    worst, bound = _check_plan(dtype, L, 3 if L <= 17 else 1, 1500 + L)
    print(f"L={L} {t} worst={worst:.3e} {sched}", flush=True)
    assert worst <= bound, (L, worst)
