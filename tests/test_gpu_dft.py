"""Parity of the GPU complex DFT with the FP64 oracle and with the reference engine (gpu)."""
from __future__ import annotations

import numpy as np
import pytest

import kfr_amd
import oracle
from tests.common import bound_fp64, fp64_dft, tol_kfr

pytestmark = pytest.mark.gpu

POW2 = [1 << k for k in range(0, 18)]
MIXED = [3, 5, 6, 7, 12, 15, 18, 30, 48, 60, 210, 1029, 3150, 8000, 15360, 30240, 7 * 4096]
GENERIC = [11, 13, 17, 31, 53, 97, 2 * 97, 9990]          # prime factors 11..97: direct DFT passes
BLUESTEIN = [101, 113, 127, 211, 3163, 96001]              # leftover factor >= 101


def _check(n, dtype, inverse, seed):
    x = oracle.complex_uniform(seed, n, dtype)
    plan = kfr_amd.dft_plan(n, dtype)
    out = np.empty_like(x)
    plan.execute(out, x, inverse=inverse)
    ref = fp64_dft(x.astype(np.complex128), inverse)
    err = oracle.rel_rms(out, ref)
    bound = bound_fp64(n, dtype)
    print(f"n={n} {np.dtype(dtype).name} inv={inverse} rel_rms_fp64={err:.3e} sched={kfr_amd.schedule('f32' if dtype == np.complex64 else 'f64', n)}",
          flush=True)
    assert err <= bound, (n, dtype, inverse, err)
    ref_engine = oracle.reference()
    if ref_engine is not None and n <= 1 << 17:
        k = ref_engine.dft(x, inverse)
        ek = oracle.rel_rms(out, k)
        assert ek <= tol_kfr(n, dtype) * 2, (n, dtype, inverse, ek)


@pytest.mark.parametrize("n", POW2 + MIXED + GENERIC + BLUESTEIN)
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_dft_sizes(n, dtype):
    _check(n, dtype, False, 1000 + n)
    _check(n, dtype, True, 2000 + n)


@pytest.mark.parametrize("n", [1 << 20, 1 << 22])
def test_dft_large_f32(n):
    _check(n, np.complex64, False, 77)


def test_dft_round_trip_2p24():
    """forward then inverse of 2^24 points returns N x (size-independent property at full size)"""
    torch = pytest.importorskip("torch")
    n = 1 << 24
    x = oracle.complex_uniform(5, n, np.complex64)
    plan = kfr_amd.dft_plan(n, np.complex64)
    xd = torch.from_numpy(x.view(np.float32)).cuda()
    y = torch.empty_like(xd)
    z = torch.empty_like(xd)
    plan.execute_batch(y, xd, 1, False)
    plan.execute_batch(z, y, 1, True)
    torch.cuda.synchronize()
    zz = z.cpu().numpy().view(np.complex64) / n
    err = oracle.rel_rms(zz, x)
    print(f"2^24 round trip rel_rms={err:.3e}", flush=True)
    assert err < 1e-6
    # spectrum against FP64 numpy
    Y = y.cpu().numpy().view(np.complex64)
    e2 = oracle.rel_rms(Y, np.fft.fft(x.astype(np.complex128)))
    print(f"2^24 forward rel_rms_fp64={e2:.3e}", flush=True)
    assert e2 < 1e-6


@pytest.mark.parametrize("n,batch", [(16, 1000), (64, 333), (1000, 77), (4096, 64), (8192, 9), (65536, 5), (101, 50)])
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_dft_batch_device(n, batch, dtype):
    torch = pytest.importorskip("torch")
    x = oracle.complex_uniform(31 + n, (batch, n), dtype)
    rdt = np.float32 if dtype == np.complex64 else np.float64
    xd = torch.from_numpy(x.view(rdt).copy()).cuda()
    yd = torch.empty_like(xd)
    plan = kfr_amd.dft_plan(n, dtype)
    for inverse in (False, True):
        plan.execute_batch(yd, xd, batch, inverse)
        torch.cuda.synchronize()
        y = yd.cpu().numpy().view(dtype).reshape(batch, n)
        ref = fp64_dft(x.astype(np.complex128), inverse)
        worst = max(oracle.rel_rms(y[b], ref[b]) for b in range(batch))
        assert worst <= bound_fp64(n, dtype), (n, batch, inverse, worst)
    # in place
    plan.execute_batch(xd, xd, batch, False)
    torch.cuda.synchronize()
    y = xd.cpu().numpy().view(dtype).reshape(batch, n)
    assert oracle.rel_rms(y, fp64_dft(x.astype(np.complex128))) <= bound_fp64(n, dtype)


def test_capi_round_trip_known_answer():
    """tests/capi_test.cpp:52-90: x[i] = (i/32, -i/32), forward then inverse gives (i, -i)"""
    for dtype in (np.complex64, np.complex128):
        n = 32
        x = np.array([complex(i / 32, -i / 32) for i in range(n)], dtype=dtype)
        p = kfr_amd.dft_plan(n, dtype)
        y = np.empty_like(x)
        z = np.empty_like(x)
        p.execute(y, x)
        p.execute(z, y, inverse=True)
        expect = np.array([complex(i, -i) for i in range(n)])
        assert np.max(np.abs(z - expect)) < 1e-5


def test_in_place_host():
    n = 1000
    x = oracle.complex_uniform(3, n, np.complex64)
    y = x.copy()
    p = kfr_amd.dft_plan(n, np.complex64)
    p.execute(y, y)
    assert oracle.rel_rms(y, fp64_dft(x.astype(np.complex128))) < 1e-6


def test_reconstruction_sine():
    """tests/dft_test.cpp:181-221: forward + inverse / N of sin(2 pi i / 96)"""
    for n in list(range(2, 257, 7)) + [9988, 9999, 96001]:
        x = np.sin(2 * np.pi * np.arange(n) / 96).astype(np.complex64)
        p = kfr_amd.dft_plan(n, np.complex64)
        y = np.empty_like(x)
        p.execute(y, x)
        p.execute(y, y, inverse=True)
        y /= n
        d = np.abs(y - x)
        assert np.sqrt(np.mean(d ** 2)) < 1e-6 and d.max() < 1e-5, n


@pytest.mark.parametrize("n", [30240, 59049, 78125, 3 * 5 ** 6 * 7])
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_generic_four_step_batch(n, dtype):
    """lengths above one workgroup with factors 3, 5, 7: four-step over the generic kernels with odd
    column and row counts, so workgroups straddle batch items and the last one is partial"""
    torch = pytest.importorskip("torch")
    batch = 3
    x = oracle.complex_uniform(700 + n % 97, (batch, n), dtype)
    rdt = np.float32 if dtype == np.complex64 else np.float64
    xd = torch.from_numpy(x.view(rdt).copy()).cuda()
    yd = torch.empty_like(xd)
    plan = kfr_amd.dft_plan(n, dtype)
    sched = kfr_amd.schedule("f32" if dtype == np.complex64 else "f64", n)
    assert "four-step" in sched, sched
    bound = bound_fp64(n, dtype)
    for inverse in (False, True):
        ref = fp64_dft(x.astype(np.complex128), inverse)
        plan.execute_batch(yd, xd, batch, inverse)
        torch.cuda.synchronize()
        y = yd.cpu().numpy().view(dtype).reshape(batch, n)
        errs = [oracle.rel_rms(y[b], ref[b]) for b in range(batch)]
        print(f"n={n} {np.dtype(dtype).name} inv={inverse} worst={max(errs):.3e} {sched}", flush=True)
        assert max(errs) <= bound, (n, inverse, errs)


@pytest.mark.parametrize("n", [96001, 100003])
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_bluestein_chain_batch(n, dtype):
    """Bluestein whose M-point transforms run as separate launches (M = 2^18): batched, both
    directions, against the FP64 DFT"""
    torch = pytest.importorskip("torch")
    batch = 3
    x = oracle.complex_uniform(800 + n % 89, (batch, n), dtype)
    rdt = np.float32 if dtype == np.complex64 else np.float64
    xd = torch.from_numpy(x.view(rdt).copy()).cuda()
    yd = torch.empty_like(xd)
    plan = kfr_amd.dft_plan(n, dtype)
    sched = kfr_amd.schedule("f32" if dtype == np.complex64 else "f64", n)
    assert "bluestein " in sched and "m=262144" in sched, sched
    bound = bound_fp64(n, dtype)
    for inverse in (False, True):
        ref = fp64_dft(x.astype(np.complex128), inverse)
        plan.execute_batch(yd, xd, batch, inverse)
        torch.cuda.synchronize()
        y = yd.cpu().numpy().view(dtype).reshape(batch, n)
        errs = [oracle.rel_rms(y[b], ref[b]) for b in range(batch)]
        print(f"n={n} {np.dtype(dtype).name} inv={inverse} worst={max(errs):.3e} {sched}", flush=True)
        assert max(errs) <= bound, (n, inverse, errs)
