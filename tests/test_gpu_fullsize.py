"""Full-size parity: the real-to-complex element of the plain sweep (f32 N = 2^24, batch 16), the real
repack over several rounds with a partial last round, the mixed-length batch workload at its full size
(20,000 transforms, every one checked), plus the host-side contracts around them: synchronous execute
after device producers on another stream, one plan used from two streams, and allocation failures
reported through kfr_last_error (gpu)."""
from __future__ import annotations

import numpy as np
import pytest

import kfr_amd
import oracle
from tests.common import bound_fp64, fp64_rdft, uneven_lengths

pytestmark = pytest.mark.gpu


def test_r2c_f32_2p24_x16():
    """dft_plan_real<float>(2^24) over a batch of 16 standard Gaussian signals (1 GiB): items 0, 5 and 15
    against the FP64 real DFT, every item through the C2R round trip"""
    torch = pytest.importorskip("torch")
    n, batch = 1 << 24, 16
    g = torch.Generator(device="cuda")
    g.manual_seed(2024)
    x = torch.randn(batch, n, device="cuda", generator=g)
    p = kfr_amd.dft_plan_real(n, kfr_amd.CCs, np.float32)
    cs = p.complex_size()
    assert cs == n // 2 + 1
    X = torch.empty(batch, 2 * cs, device="cuda")
    p.execute_batch(X, x, batch, False)
    torch.cuda.synchronize()
    bound = bound_fp64(n, np.float32)
    for b in (0, 5, batch - 1):
        ref = np.fft.rfft(x[b].cpu().numpy().astype(np.float64))
        e = oracle.rel_rms(X[b].cpu().numpy().view(np.complex64), ref)
        print(f"R2C 2^24 x16 item {b}: rel_rms_fp64={e:.3e}", flush=True)
        assert e <= bound, (b, e)
    y = torch.empty_like(x)
    p.execute_batch(y, X, batch, True)
    torch.cuda.synchronize()
    worst = 0.0
    for b in range(batch):
        xb = x[b].double()
        worst = max(worst, float((y[b].double() / n - xb).norm() / xb.norm()))
    print(f"R2C 2^24 x16 round trip worst rel_rms={worst:.3e}", flush=True)
    assert worst <= bound


@pytest.mark.parametrize("fmt", [kfr_amd.Perm, kfr_amd.CCs])
def test_r2c_repack_several_rounds(fmt, monkeypatch):
    """half spectra staged in rounds of 48 MiB: 2^22-point items (16 MiB of half spectrum each) go three
    to a round, so a batch of 7 runs rounds of 3, 3 and 1; every item against the FP64 real DFT"""
    torch = pytest.importorskip("torch")
    monkeypatch.setenv("KFR_AMD_REAL_CHUNK_MB", "48")
    n, batch = 1 << 22, 7
    x = oracle.gaussian(4242 + fmt, batch * n, np.float32).reshape(batch, n)
    xd = torch.from_numpy(x).cuda()
    p = kfr_amd.dft_plan_real(n, fmt, np.float32)
    cs = p.complex_size()
    X = torch.full((batch, 2 * cs), float("nan"), device="cuda")
    p.execute_batch(X, xd, batch, False)
    torch.cuda.synchronize()
    got = X.cpu().numpy().view(np.complex64)
    ref = fp64_rdft(x.astype(np.float64), fmt)
    errs = [oracle.rel_rms(got[b], ref[b]) for b in range(batch)]
    print(f"R2C 2^22 x7 fmt={fmt} in rounds of 3: worst rel_rms={max(errs):.3e}", flush=True)
    assert max(errs) <= bound_fp64(n, np.float32), errs
    y = torch.empty_like(xd)
    p.execute_batch(y, X, batch, True)
    torch.cuda.synchronize()
    e2 = oracle.rel_rms(y.cpu().numpy() / n, x)
    assert e2 <= bound_fp64(n, np.float32), e2


def test_uneven_batch_full_workload():
    """the mixed-length workload at full size: 20,000 complex-f32 transforms with lengths uniform over
    the 328 7-smooth N in [18, 8192] and primes in [17, 127], directions Bernoulli(1/2), data U[-1, 1)
    (seed 3, the same draw as the benchmark), one autotuned persistent launch; every transform
    against the FP64 DFT"""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    L = uneven_lengths()
    sizes = rng.choice(L, 20000)
    inv = rng.integers(0, 2, sizes.size).astype(np.int32)
    total = int(sizes.sum())
    x = (rng.uniform(-1, 1, 2 * total).astype(np.float32)).view(np.complex64)
    b = kfr_amd.dft_batch(sizes, inv, np.complex64, autotune=True)
    xd = torch.from_numpy(x.view(np.float32).copy()).cuda()
    yd = torch.full_like(xd, float("nan"))
    b.execute(yd, xd)
    torch.cuda.synchronize()
    y = yd.cpu().numpy().view(np.complex64)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    groups: dict = {}
    for i, (n, d) in enumerate(zip(sizes, inv)):
        groups.setdefault((int(n), int(d)), []).append(i)
    worst, where = 0.0, None
    for (n, d), idx in groups.items():
        xs = np.stack([x[offs[i]:offs[i] + n] for i in idx]).astype(np.complex128)
        ys = np.stack([y[offs[i]:offs[i] + n] for i in idx]).astype(np.complex128)
        ref = np.fft.ifft(xs, axis=1) * n if d else np.fft.fft(xs, axis=1)
        e = np.linalg.norm(ys - ref, axis=1) / np.linalg.norm(ref, axis=1)
        if e.max() > worst:
            worst, where = float(e.max()), (n, d)
    print(f"uneven batch: {sizes.size} transforms, {total} points, {len(groups)} (length, direction) groups, "
          f"worst rel_rms={worst:.3e} at {where}", flush=True)
    assert worst <= 1e-6, (worst, where)


def test_sync_execute_waits_for_producers_on_other_streams():
    """the synchronous execute of a device buffer that another stream is still writing: the transform
    sees the final values"""
    torch = pytest.importorskip("torch")
    n = 1 << 20
    x = oracle.complex_uniform(61, n, np.complex64)
    src = torch.from_numpy(x.view(np.float32).copy()).cuda()
    xd = torch.zeros_like(src)
    yd = torch.empty_like(src)
    p = kfr_amd.dft_plan(n, np.complex64)
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        junk = torch.ones(1 << 26, device="cuda")
        for _ in range(40):  # keep the side stream busy before it writes the input
            junk.mul_(1.0000001)
        xd.copy_(src)
    p.execute(yd, xd)  # no synchronisation in between
    torch.cuda.synchronize()
    e = oracle.rel_rms(yd.cpu().numpy().view(np.complex64), np.fft.fft(x.astype(np.complex128)))
    assert e <= bound_fp64(n, np.complex64), e


def test_one_plan_on_two_streams():
    """a four-step plan (one shared intermediate) used on two streams back to back: the second use waits
    for the first, both results are right"""
    torch = pytest.importorskip("torch")
    n, batch = 1 << 20, 24
    x = oracle.complex_uniform(62, (batch, n), np.complex64)
    xd = torch.from_numpy(x.view(np.float32).copy()).cuda()
    ya = torch.empty_like(xd)
    yb = torch.empty_like(xd)
    p = kfr_amd.dft_plan(n, np.complex64)
    assert "four-step" in kfr_amd.schedule("f32", n)
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    p.execute_batch(ya, xd, batch, False, sa)
    p.execute_batch(yb, xd, batch, True, sb)
    torch.cuda.synchronize()
    xx = x.astype(np.complex128)
    fa = np.fft.fft(xx[[0, batch - 1]], axis=1)
    fb = np.fft.ifft(xx[[0, batch - 1]], axis=1) * n
    a = ya.cpu().numpy().view(np.complex64).reshape(batch, n)[[0, batch - 1]]
    b = yb.cpu().numpy().view(np.complex64).reshape(batch, n)[[0, batch - 1]]
    assert oracle.rel_rms(a, fa) <= 1e-6 and oracle.rel_rms(b, fb) <= 1e-6


def test_injected_allocation_failure_reaches_last_error():
    """a device allocation failure (forced from the host) leaves NULL from the creator and the HIP
    error text in kfr_last_error; the next calls work and clear it"""
    L = kfr_amd.lib()
    L.kfr_amd_inject_alloc_failure(1)
    try:
        h = L.kfr_dft_create_plan_f32(4096)
        assert not h
        assert "injected allocation failure" in kfr_amd.last_error()
    finally:
        L.kfr_amd_inject_alloc_failure(0)
    h = L.kfr_dft_create_plan_f32(4096)
    assert h and kfr_amd.last_error() == ""
    L.kfr_dft_delete_plan_f32(h)
    # an execute whose staging buffer cannot be allocated reports it and leaves the plan usable
    p = kfr_amd.dft_plan(8192, np.complex64)
    x = oracle.complex_uniform(63, 8192, np.complex64)
    y = np.empty_like(x)
    L.kfr_amd_inject_alloc_failure(1)
    try:
        with pytest.raises(kfr_amd.KfrError, match="injected allocation failure"):
            p.execute(y, x)
    finally:
        L.kfr_amd_inject_alloc_failure(0)
    p.execute(y, x)
    assert oracle.rel_rms(y, np.fft.fft(x.astype(np.complex128))) <= 1e-6
