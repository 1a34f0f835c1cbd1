"""Parity of the GPU real DFT, DCT, convolution, streaming filter and convolution adjoint (gpu)."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import kfr_amd
import oracle
from tests.common import bound_fp64, fp64_dct, fp64_rdft, tol_kfr

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
KAT = json.loads((GOLD / "kat.json").read_text())
SP = np.load(GOLD / "spectra.npz")


@pytest.mark.parametrize("n", [2, 4, 6, 16, 30, 64, 202, 2048, 4096, 30720, 1 << 15, 1 << 17, 6326, 1 << 20])
@pytest.mark.parametrize("fmt", [0, 1])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_real_dft(n, fmt, dt):
    x = oracle.gaussian(700 + n, n, dt)
    p = kfr_amd.dft_plan_real(n, fmt, dt)
    cdt = np.complex64 if dt == np.float32 else np.complex128
    X = np.empty(p.complex_size(), dtype=cdt)
    p.execute(X, x)
    ref = fp64_rdft(x.astype(np.float64), fmt)
    e = oracle.rel_rms(X, ref)
    bound = bound_fp64(n, dt)
    print(f"real n={n} fmt={fmt} {np.dtype(dt).name} fwd rel_rms={e:.3e}", flush=True)
    assert e <= bound
    y = np.empty(n, dtype=dt)
    p.execute(y, X)
    e2 = oracle.rel_rms(y / n, x)
    assert e2 <= bound * 2, e2


@pytest.mark.parametrize("n,dt", [(512, np.float32), (2048, np.float32), (1 << 14, np.float32), (1 << 16, np.float32),
                                  (256, np.float64), (2048, np.float64), (1 << 15, np.float64)])
@pytest.mark.parametrize("fmt", [0, 1])
def test_real_one_kernel_ragged_batch(n, dt, fmt):
    """R2C and C2R in one kernel (power-of-two halves 2^8..2^15 f32, 2^7..2^14 f64) over a batch that
    leaves the last workgroup part-filled; every transform against FP64"""
    torch = pytest.importorskip("torch")
    b = 37
    cdt = np.complex64 if dt == np.float32 else np.complex128
    x = oracle.gaussian(88 + n + fmt, n * b, dt).reshape(b, n)
    p = kfr_amd.dft_plan_real(n, fmt, dt)
    cs = p.complex_size()
    xd = torch.from_numpy(x).cuda()
    Xd = torch.full((b, cs * 2), float("nan"), dtype=xd.dtype, device="cuda")
    p.execute_batch(Xd, xd, b)
    torch.cuda.synchronize()
    X = Xd.cpu().numpy().view(cdt).reshape(b, cs)
    ref = fp64_rdft(x.astype(np.float64), fmt)
    bound = 1e-6 if dt == np.float32 else 1e-14
    assert max(oracle.rel_rms(X[i], ref[i]) for i in range(b)) <= bound
    yd = torch.full_like(xd, float("nan"))
    p.execute_batch(yd, torch.from_numpy(ref.astype(cdt).view(dt)).cuda(), b, True)
    torch.cuda.synchronize()
    y = yd.cpu().numpy().reshape(b, n)
    assert max(oracle.rel_rms(y[i] / n, x[i]) for i in range(b)) <= bound


@pytest.mark.parametrize("tn", ["f32", "f64"])
def test_real_dft_matches_reference_outputs(tn):
    dt = np.float32 if tn == "f32" else np.float64
    for n in [2, 4, 30, 64, 2048]:
        x = SP[f"r_in_{tn}_{n}"]
        for fmt in (0, 1):
            p = kfr_amd.dft_plan_real(n, fmt, dt)
            X = np.empty(p.complex_size(), dtype=SP[f"r_fwd_{tn}_{n}_{fmt}"].dtype)
            p.execute(X, x)
            assert oracle.rel_rms(X, SP[f"r_fwd_{tn}_{n}_{fmt}"]) <= 2 * tol_kfr(n, dt)
            y = np.empty(n, dtype=dt)
            p.execute(y, SP[f"r_fwd_{tn}_{n}_{fmt}"])
            assert oracle.rel_rms(y, SP[f"r_inv_{tn}_{n}_{fmt}"]) <= 2 * tol_kfr(n, dt)


def test_dct_known_answers():
    p = kfr_amd.dct_plan(16, np.float32)
    x = np.arange(16, dtype=np.float32)
    y = np.empty(16, np.float32)
    p.execute(y, x)
    assert np.sqrt(np.mean((y - KAT["dct_fwd"]) ** 2)) < 1e-5
    p.execute(y, x, inverse=True)
    assert np.sqrt(np.mean((y - KAT["dct_inv"]) ** 2)) < 1e-5


@pytest.mark.parametrize("n", [1, 2, 3, 8, 12, 16, 17, 64, 256, 512, 1000, 4096, 7680, 8000, 32768, 65536, 1 << 20])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_dct(n, dt):
    """lengths 4m whose half has a power-of-two or menu plan take the half-length real-transform path
    (8, 16, 64, 4096, 7680, 8000, 65536, 2^20), the others the n-point generic path"""
    x = oracle.gaussian(900 + n, n, dt)
    p = kfr_amd.dct_plan(n, dt)
    for inv in (False, True):
        y = np.empty(n, dtype=dt)
        p.execute(y, x, inverse=inv)
        ref = fp64_dct(x.astype(np.float64), inv)
        e = oracle.rel_rms(y, ref)
        assert e <= max(2e-6 if dt == np.float32 else 1e-13, tol_kfr(n, dt)), (n, inv, e)


@pytest.mark.parametrize("n,dt", [(512, np.float32), (1024, np.float32), (2048, np.float32), (8192, np.float32),
                                  (16384, np.float32), (32768, np.float32), (65536, np.float32), (256, np.float64),
                                  (512, np.float64), (1024, np.float64), (2048, np.float64), (8192, np.float64),
                                  (16384, np.float64), (32768, np.float64)])
def test_dct2_one_kernel_ragged_batch(n, dt):
    """the one-kernel DCT-II and DCT-III (half lengths 2^8..2^15 in f32, 2^7..2^14 in f64: with and
    without the mirrored first and last passes) over a batch that leaves the last workgroup
    part-filled; every transform against FP64"""
    torch = pytest.importorskip("torch")
    b = 37
    x = oracle.gaussian(77 + n, n * b, dt)
    xd = torch.from_numpy(x).cuda()
    p = kfr_amd.dct_plan(n, dt)
    for inverse in (False, True):
        yd = torch.full_like(xd, float("nan"))
        p.execute_batch(yd, xd, b, inverse)
        torch.cuda.synchronize()
        y = yd.cpu().numpy().reshape(b, n)
        ref = fp64_dct(x.reshape(b, n).astype(np.float64), inverse)
        e = max(oracle.rel_rms(y[i], ref[i]) for i in range(b))
        assert e <= (1e-6 if dt == np.float32 else 1e-14), (inverse, e)


def test_dct_batch_65536_x_1024_sample():
    torch = pytest.importorskip("torch")
    n, b = 65536, 64
    x = oracle.gaussian(42, n * b, np.float32)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    p = kfr_amd.dct_plan(n, np.float32)
    p.execute_batch(yd, xd, b)
    torch.cuda.synchronize()
    y = yd.cpu().numpy().reshape(b, n)
    ref = fp64_dct(x.reshape(b, n)[:4].astype(np.float64))
    assert max(oracle.rel_rms(y[i], ref[i]) for i in range(4)) < 2e-6
    p.execute_batch(yd, xd, b, True)
    torch.cuda.synchronize()
    y = yd.cpu().numpy().reshape(b, n)
    ref = fp64_dct(x.reshape(b, n)[-4:].astype(np.float64), True)
    assert max(oracle.rel_rms(y[b - 4 + i], ref[i]) for i in range(4)) < 2e-6


def test_convolution_known_answers():
    for dt in (np.float32, np.float64):
        a = np.array(KAT["conv_a"], dtype=dt)
        b = np.array(KAT["conv_b"], dtype=dt)
        assert np.sqrt(np.mean((kfr_amd.convolve(a, b) - KAT["convolve"]) ** 2)) < 1e-4
        assert np.sqrt(np.mean((kfr_amd.correlate(a, b) - KAT["correlate"]) ** 2)) < 1e-4
    ac = np.array(KAT["conv_a"], dtype=np.complex64)
    bc = np.array(KAT["conv_b"], dtype=np.complex64)
    assert np.sqrt(np.mean(np.abs(kfr_amd.convolve(ac, bc) - KAT["convolve"]) ** 2)) < 1e-4


@pytest.mark.parametrize("tn", ["f32", "f64"])
def test_convolution_matches_reference_outputs(tn):
    a, b = SP[f"conv_a_{tn}"], SP[f"conv_b_{tn}"]
    tol = 1e-5 if tn == "f32" else 1e-12
    assert oracle.rel_rms(kfr_amd.convolve(a, b), SP[f"conv_{tn}"]) < tol
    assert oracle.rel_rms(kfr_amd.correlate(a, b), SP[f"corr_{tn}"]) < tol
    assert oracle.rel_rms(kfr_amd.autocorrelate(a), SP[f"acorr_{tn}"]) < tol
    out = kfr_amd.convolve(SP["cconv_a"], SP["cconv_b"])
    assert np.allclose(out.imag, 0) and oracle.rel_rms(out, SP["cconv"]) < 1e-5


def test_convolve_filter_known_answers_and_state():
    for dt in (np.float32, np.float64):
        a = np.array(KAT["conv_a"], dtype=dt)
        b = np.array(KAT["conv_b"], dtype=dt)
        f = kfr_amd.convolve_filter(a)
        dest = np.zeros(5, dtype=dt)
        f.apply(dest, b)
        assert np.sqrt(np.mean((dest - KAT["convolve_filter_first5"]) ** 2)) < 1e-4
        f.apply(dest, b)   # the state carries over: the second block differs
        assert np.sqrt(np.mean((dest - KAT["convolve_filter_first5"]) ** 2)) > 1e-4
        f.reset()
        f.apply(dest, b)
        assert np.sqrt(np.mean((dest - KAT["convolve_filter_first5"]) ** 2)) < 1e-4


@pytest.mark.parametrize("tn", ["f32", "f64"])
def test_convolve_filter_streaming_matches_reference(tn):
    dt = np.float32 if tn == "f32" else np.float64
    taps, sig = SP[f"filt_taps_{tn}"], SP[f"filt_in_{tn}"]
    f = kfr_amd.convolve_filter(taps, 64)
    chunks = [sig[:1], sig[1:700], sig[700:701], sig[701:3000], sig[3000:]]
    outs = []
    for c in chunks:
        o = np.empty_like(c)
        f.apply(o, c.copy())
        outs.append(o)
    out = np.concatenate(outs)
    assert oracle.rel_rms(out, SP[f"filt_out_{tn}"]) < (2e-5 if tn == "f32" else 1e-12)


def test_convolve_filter_4095_taps_block_1024():
    """the convolution filter of the C API with 4095 taps and block 1024 (Hann-windowed sinc)"""
    n = 4095
    t = np.arange(n) - (n - 1) / 2
    h = (0.4 * np.sinc(0.4 * t) * np.hanning(n)).astype(np.float32)
    x = oracle.gaussian(3, 20000, np.float32)
    f = kfr_amd.convolve_filter(h, 1024)
    y = np.empty_like(x)
    f.apply(y[:5000], x[:5000].copy())
    f.apply(y[5000:], x[5000:].copy())
    ref = oracle.restated.convolve(x.astype(np.float64), h.astype(np.float64))[: x.size]
    assert oracle.rel_rms(y, ref) < 1e-5


def test_batched_conv_forward_and_adjoint_fd_fp64():
    torch = pytest.importorskip("torch")
    from kfr_amd.autograd import fft_convolve
    torch.manual_seed(0)
    nx, nh, B = 4096, 255, 2
    x = torch.from_numpy(oracle.gaussian(1, B * nx, np.float64).reshape(B, nx)).cuda().requires_grad_(True)
    h = torch.from_numpy(oracle.gaussian(2, B * nh, np.float64).reshape(B, nh)).cuda().requires_grad_(True)
    y = fft_convolve(x, h)
    yr = np.stack([np.convolve(x[b].detach().cpu().numpy(), h[b].detach().cpu().numpy()) for b in range(B)])
    assert oracle.rel_rms(y.detach().cpu().numpy(), yr) < 1e-13
    w = torch.from_numpy(oracle.gaussian(3, B * (nx + nh - 1), np.float64).reshape(B, -1)).cuda()
    (y * w).sum().backward()
    # finite differences at 64 random coordinates (central, FP64)
    rng = np.random.default_rng(7)

    def loss(xv, hv):
        return float((fft_convolve(xv, hv) * w).sum())

    eps = 1e-3
    worst = 0.0
    for _ in range(32):
        b, i = int(rng.integers(B)), int(rng.integers(nx))
        xp = x.detach().clone()
        xm = x.detach().clone()
        xp[b, i] += eps
        xm[b, i] -= eps
        fd = (loss(xp, h.detach()) - loss(xm, h.detach())) / (2 * eps)
        worst = max(worst, abs(fd - float(x.grad[b, i])))
    for _ in range(32):
        b, j = int(rng.integers(B)), int(rng.integers(nh))
        hp = h.detach().clone()
        hm = h.detach().clone()
        hp[b, j] += eps
        hm[b, j] -= eps
        fd = (loss(x.detach(), hp) - loss(x.detach(), hm)) / (2 * eps)
        worst = max(worst, abs(fd - float(h.grad[b, j])))
    print(f"adjoint vs central differences (FP64, 64 coordinates): max abs diff {worst:.3e}", flush=True)
    assert worst < 1e-8


@pytest.mark.parametrize("nx,nh", [(4096, 255), (300, 200)])
def test_batched_correlate_adjoint_fd_fp64(nx, nh):
    """the correlation (convolution with h reversed) and its adjoint, on the overlap-save (4096 x 255)
    and the full-length (300 x 200) paths: forward against numpy, gradients against FP64 central
    differences on 64 coordinates (32 of the signal, 32 of the taps)"""
    torch = pytest.importorskip("torch")
    from kfr_amd.autograd import fft_correlate
    B = 2
    x = torch.from_numpy(oracle.gaussian(21, B * nx, np.float64).reshape(B, nx)).cuda().requires_grad_(True)
    h = torch.from_numpy(oracle.gaussian(22, B * nh, np.float64).reshape(B, nh)).cuda().requires_grad_(True)
    y = fft_correlate(x, h)
    yr = np.stack([np.convolve(x[b].detach().cpu().numpy(), h[b].detach().cpu().numpy()[::-1]) for b in range(B)])
    assert oracle.rel_rms(y.detach().cpu().numpy(), yr) < 1e-13
    w = torch.from_numpy(oracle.gaussian(23, B * (nx + nh - 1), np.float64).reshape(B, -1)).cuda()
    (y * w).sum().backward()
    rng = np.random.default_rng(8)

    def loss(xv, hv):
        return float((fft_correlate(xv, hv) * w).sum())

    eps = 1e-3
    worst = 0.0
    for k in range(64):
        on_x = k < 32
        b = int(rng.integers(B))
        i = int(rng.integers(nx if on_x else nh))
        v = (x if on_x else h).detach()
        vp, vm = v.clone(), v.clone()
        vp[b, i] += eps
        vm[b, i] -= eps
        if on_x:
            fd = (loss(vp, h.detach()) - loss(vm, h.detach())) / (2 * eps)
            worst = max(worst, abs(fd - float(x.grad[b, i])))
        else:
            fd = (loss(x.detach(), vp) - loss(x.detach(), vm)) / (2 * eps)
            worst = max(worst, abs(fd - float(h.grad[b, i])))
    print(f"correlate adjoint nx={nx} nh={nh} vs central differences: max abs diff {worst:.3e}", flush=True)
    assert worst < 1e-8


def test_batched_conv_adjoint_matches_torch_fft_reference():
    torch = pytest.importorskip("torch")
    from kfr_amd.autograd import fft_convolve
    nx, nh, B = 1 << 16, 4095, 4
    x = torch.from_numpy(oracle.gaussian(11, B * nx, np.float32).reshape(B, nx)).cuda().requires_grad_(True)
    h = torch.from_numpy(oracle.gaussian(12, nh, np.float32)).cuda().requires_grad_(True)
    y = fft_convolve(x, h)
    g = torch.from_numpy(oracle.gaussian(13, B * (nx + nh - 1), np.float32).reshape(B, -1)).cuda()
    y.backward(g)
    # torch.fft reference gradient in float64 on the CPU
    xr = x.detach().double().cpu().requires_grad_(True)
    hr = h.detach().double().cpu().requires_grad_(True)
    S = 1 << 17
    yr = torch.fft.irfft(torch.fft.rfft(xr, S) * torch.fft.rfft(hr, S), S)[:, : nx + nh - 1]
    yr.backward(g.double().cpu())
    assert oracle.rel_rms(y.detach().cpu().numpy(), yr.detach().numpy()) < 1e-5
    assert oracle.rel_rms(x.grad.cpu().numpy(), xr.grad.numpy()) < 1e-5
    assert oracle.rel_rms(h.grad.cpu().numpy(), hr.grad.numpy()) < 1e-5


def test_conv_full_size_per_signal_taps():
    """the full-size convolution workload: 128 signals of 2^20 samples, each with its own 4095-tap
    Hann-windowed sinc (cutoff 0.2 fs, plus 1e-3 noise); output and adjoint on three sampled signals
    against torch.fft (rocFFT) in float64"""
    torch = pytest.importorskip("torch")
    from kfr_amd.autograd import fft_convolve
    B, nx, nh = 128, 1 << 20, 4095
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    x = torch.randn(B, nx, device="cuda", generator=g).requires_grad_(True)
    m = torch.arange(nh, device="cuda", dtype=torch.float64) - (nh - 1) / 2
    base = 0.4 * torch.sinc(0.4 * m) * torch.hann_window(nh, periodic=False, device="cuda", dtype=torch.float64)
    h = (base.float()[None, :] + 1e-3 * torch.randn(B, nh, device="cuda", generator=g)).requires_grad_(True)
    y = fft_convolve(x, h)
    gout = torch.randn(B, nx + nh - 1, device="cuda", generator=g)
    y.backward(gout)
    torch.cuda.synchronize()
    pick = [0, 77, 127]
    S = 1 << 21
    xr = x.detach()[pick].double().requires_grad_(True)
    hr = h.detach()[pick].double().requires_grad_(True)
    yr = torch.fft.irfft(torch.fft.rfft(xr, S) * torch.fft.rfft(hr, S), S)[:, : nx + nh - 1]
    yr.backward(gout[pick].double())

    def rel(a, b):
        return float((a.detach().double() - b.detach()).norm() / b.detach().norm())

    ey, ex, eh = rel(y[pick], yr), rel(x.grad[pick], xr.grad), rel(h.grad[pick], hr.grad)
    print(f"full-size conv rel_rms y={ey:.3e} dx={ex:.3e} dh={eh:.3e}", flush=True)
    assert ey < 1e-6 and ex < 1e-6 and eh < 1e-6


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("nx,nh", [(100000, 3), (50001, 257), (70000, 1500), (1 << 15, 2049), (200003, 4001)])
def test_overlap_save_lengths(dt, nx, nh):
    """long signals with short filters take the overlap-save kernels: odd segment counts, a last
    segment pair with one segment, filters up to half the block; convolve and correlate against
    numpy in FP64"""
    a = oracle.gaussian(31 + nh, nx, dt)
    b = oracle.gaussian(37 + nh, nh, dt)
    tol = 1e-5 if dt == np.float32 else 1e-12
    ref = np.convolve(a.astype(np.float64), b.astype(np.float64))
    e = oracle.rel_rms(kfr_amd.convolve(a, b), ref)
    refc = np.convolve(a.astype(np.float64), b[::-1].astype(np.float64))
    ec = oracle.rel_rms(kfr_amd.correlate(a, b), refc)
    print(f"overlap-save {np.dtype(dt).name} nx={nx} nh={nh} conv={e:.3e} corr={ec:.3e}", flush=True)
    assert e < tol and ec < tol


def test_conv_shared_filter_gradient_many_items():
    """a filter shared by 96 signals: its gradient sums 96 x (segments) partial correlations"""
    torch = pytest.importorskip("torch")
    from kfr_amd.autograd import fft_convolve
    B, nx, nh = 96, 1 << 16, 1000
    g = torch.Generator(device="cuda")
    g.manual_seed(9)
    x = torch.randn(B, nx, device="cuda", generator=g).requires_grad_(True)
    h = torch.randn(nh, device="cuda", generator=g).requires_grad_(True)
    y = fft_convolve(x, h)
    gout = torch.randn(B, nx + nh - 1, device="cuda", generator=g)
    y.backward(gout)
    S = 1 << 17
    xr = x.detach().double().requires_grad_(True)
    hr = h.detach().double().requires_grad_(True)
    yr = torch.fft.irfft(torch.fft.rfft(xr, S) * torch.fft.rfft(hr, S), S)[:, : nx + nh - 1]
    yr.backward(gout.double())

    def rel(a, b):
        return float((a.detach().double() - b.detach()).norm() / b.detach().norm())

    ey, ex, eh = rel(y, yr), rel(x.grad, xr.grad), rel(h.grad, hr.grad)
    print(f"shared filter B={B}: rel_rms y={ey:.3e} dx={ex:.3e} dh={eh:.3e}", flush=True)
    assert ey < 1e-6 and ex < 1e-6 and eh < 1e-5


def test_conv_empty_batch():
    """no signals: nothing is written by the forward call, the adjoint of a shared filter is zero"""
    torch = pytest.importorskip("torch")
    L = kfr_amd.lib()
    for nx, nh in ((1 << 16, 255), (64, 16)):  # overlap-save and full-length paths
        c = L.kfr_amd_conv_create(0, nx, nh)
        assert c
        y = torch.full((8,), 3.0, device="cuda")
        dh = torch.full((nh,), 5.0, device="cuda")
        x = torch.zeros(8, device="cuda")
        h = torch.zeros(nh, device="cuda")
        assert L.kfr_amd_conv_forward(c, y.data_ptr(), x.data_ptr(), h.data_ptr(), nh, 0, 0, None) == 0
        assert L.kfr_amd_conv_backward(c, None, dh.data_ptr(), x.data_ptr(), x.data_ptr(), h.data_ptr(), 0, 0, 0,
                                       None) == 0
        torch.cuda.synchronize()
        assert torch.all(y == 3.0) and torch.all(dh == 0.0)
        L.kfr_amd_conv_delete(c)
