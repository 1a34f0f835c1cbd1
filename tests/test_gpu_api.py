"""The class API around the transforms (gpu): the C++ class-API program over include/kfr_amd/dft.hpp,
the plan cache and the convolutions that use it, progressive (stage-at-a-time) execution, and the
complex streaming filter."""
from __future__ import annotations

import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import kfr_amd
import oracle
from tests.common import bound_fp64, fp64_dft, fp64_rdft

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
KAT = json.loads((ROOT / "tests" / "golden" / "kat.json").read_text())


def test_cpp_class_api_program():
    """tests/cpp/dft_api_test.cpp, written against kfr::dft_plan / dft_plan_real / dct_plan / dft_plan_md /
    convolve / convolve_filter / dft_cache, checks itself against the reference's known answers and the
    oracle's long-double DFT"""
    exe = ROOT / "tests" / "cpp" / "dft_api_test"
    assert exe.exists(), "tests/cpp/dft_api_test is built by __graft_entry__.build()"
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    print(r.stdout[-8000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert r.stdout.count("PASS") >= 100 and "FAIL" not in r.stdout


def test_cache_shares_plans_and_convolve_builds_no_tables():
    """the cache returns the same GPU plan for a key; a second convolve of the same sizes launches no
    table generation"""
    c = kfr_amd.dft_cache.instance()
    c.clear()
    assert c.size() == 0
    p1, p2 = c.get(np.complex64, 1000), c.get(np.complex64, 1000)
    assert p1._h == p2._h and c.size() == 1
    assert c.get(np.complex128, 1000)._h != p1._h          # the key names the type
    r1, r2 = c.getreal(np.float32, 1000), c.getreal(np.float32, 1000, kfr_amd.Perm)
    assert r1._h != r2._h and c.size() == 4                # ... and the pack format
    a = oracle.gaussian(5, 5000, np.float32)
    b = oracle.gaussian(6, 300, np.float32)
    with kfr_amd.kernel_timing() as first:
        y1 = kfr_amd.convolve(a, b)
    with kfr_amd.kernel_timing() as second:
        y2 = kfr_amd.convolve(a, b)
    names1 = [r[0] for r in first.records]
    names2 = [r[0] for r in second.records]
    assert any(n.startswith("k_twiddles") for n in names1), names1
    assert not any(n.startswith("k_twiddles") for n in names2), names2
    assert np.array_equal(y1, y2)
    assert oracle.rel_rms(y1, np.convolve(a.astype(np.float64), b.astype(np.float64))) < 1e-5
    # the plans outlive the cache's own references
    c.clear()
    x = oracle.complex_uniform(7, 1000, np.complex64)
    y = np.empty_like(x)
    p1.execute(y, x)
    assert oracle.rel_rms(y, fp64_dft(x.astype(np.complex128))) <= 1e-6


def test_cache_helpers_dft_idft_realdft_irealdft():
    """dft / idft / realdft / irealdft (include/kfr/dft/cache.hpp:128-170) through the cache"""
    for dt, rdt in ((np.complex64, np.float32), (np.complex128, np.float64)):
        x = oracle.complex_uniform(11, 4096, dt)
        b = 1e-6 if dt == np.complex64 else 1e-13
        assert oracle.rel_rms(kfr_amd.dft(x), fp64_dft(x.astype(np.complex128))) <= b
        assert oracle.rel_rms(kfr_amd.idft(x), fp64_dft(x.astype(np.complex128), True)) <= b
        r = oracle.gaussian(12, 4096, rdt)
        R = kfr_amd.realdft(r)
        assert R.size == 2049 and oracle.rel_rms(R, fp64_rdft(r.astype(np.float64), 1)) <= b
        assert oracle.rel_rms(kfr_amd.irealdft(R) / 4096, r) <= 2 * b


PROG_CASES = [(16, False), (4096, False), (4096, True), (1 << 14, False), (1 << 16, True), (1 << 20, False),
              (210, False), (1029, False), (9990, False), (3 ** 9, False), (101, False), (96001, False)]


@pytest.mark.parametrize("n,opt", PROG_CASES)
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_progressive_matches_execute(n, opt, dtype):
    """progressive_start / progressive_step: the reference's step count, real work from the first step
    (one GPU launch or more), and after the last step the same spectrum as the FP64 DFT"""
    x = oracle.complex_uniform(100 + n % 89, n, dtype)
    p = kfr_amd.dft_plan(n, dtype, progressive_optimized=opt)
    ref = fp64_dft(x.astype(np.complex128))
    out = np.full_like(x, np.nan)
    temp = np.empty(max(1, p.temp_size), np.uint8)
    prog = p.progressive_start(False, out, x, temp)
    assert prog.total == p.progressive_total_steps()
    assert prog.total == int(kfr_amd.describe(3 if opt else 0, "f32" if dtype == np.complex64 else "f64", n)
                             .splitlines()[0].split("stages=")[1].split(",")[0])
    launches = []
    more = True
    while more:
        with kfr_amd.kernel_timing() as kt:
            more = p.progressive_step(prog)
        launches.append(len(kt.records))
    assert len(launches) == prog.total
    assert launches[0] >= 1, launches
    e = oracle.rel_rms(out, ref)
    print(f"progressive n={n} opt={opt} {np.dtype(dtype).name}: {prog.total} steps, launches {launches}, "
          f"rel_rms={e:.3e}", flush=True)
    assert e <= bound_fp64(n, dtype)
    with pytest.raises(kfr_amd.KfrError, match="already complete"):
        p.progressive_step(prog)


@pytest.mark.parametrize("n", [64, 4096, 1 << 16, 30720])
@pytest.mark.parametrize("fmt", [kfr_amd.Perm, kfr_amd.CCs])
def test_progressive_real(n, fmt):
    """dft_plan_real progressive in both directions: repack (forward, last) and unpack (inverse, first)
    are units of their own"""
    x = oracle.gaussian(300 + n, n, np.float32)
    p = kfr_amd.dft_plan_real(n, fmt, np.float32)
    X = np.full(p.complex_size(), np.nan, np.complex64)
    prog = p.progressive_start(X, x, np.empty(max(1, p.temp_size), np.uint8))
    steps = 1
    while p.progressive_step(prog):
        steps += 1
    assert steps == p.progressive_total_steps()
    assert oracle.rel_rms(X, fp64_rdft(x.astype(np.float64), fmt)) <= 1e-6
    y = np.full(n, np.nan, np.float32)
    prog = p.progressive_start(y, X, np.empty(max(1, p.temp_size), np.uint8))
    while p.progressive_step(prog):
        pass
    assert oracle.rel_rms(y / n, x) <= 2e-6


def test_progressive_with_device_buffers_and_interleaved_executes():
    """a four-step progressive execution keeps its intermediate apart: executes of the same plan between
    its steps do not disturb it"""
    torch = pytest.importorskip("torch")
    n = 1 << 20
    x = oracle.complex_uniform(401, n, np.complex64)
    xd = torch.from_numpy(x.view(np.float32).copy()).cuda()
    yd = torch.empty_like(xd)
    other = torch.empty_like(xd)
    p = kfr_amd.dft_plan(n, np.complex64)
    prog = p.progressive_start(False, yd, xd, np.empty(max(1, p.temp_size), np.uint8))
    more = True
    while more:
        more = p.progressive_step(prog)
        p.execute_batch(other, xd, 1, True)   # the same plan, its own scratch
    torch.cuda.synchronize()
    ref = np.fft.fft(x.astype(np.complex128))
    assert oracle.rel_rms(yd.cpu().numpy().view(np.complex64), ref) <= 1e-6


def test_complex_convolve_filter():
    """convolve_filter<complex<T>>: the reference's known answer (tests/dft_test.cpp:131-146), state
    across calls and reset, and a long stream in uneven chunks against the direct convolution"""
    for dt in (np.complex64, np.complex128):
        a = np.array(KAT["conv_a"], dtype=dt)
        b = np.array(KAT["conv_b"], dtype=dt)
        f = kfr_amd.convolve_filter(a)
        dest = np.zeros(5, dtype=dt)
        f.apply(dest, b)
        assert np.sqrt(np.mean(np.abs(dest - KAT["convolve_filter_first5"]) ** 2)) < 1e-4
        f.apply(dest, b)
        assert np.sqrt(np.mean(np.abs(dest - KAT["convolve_filter_first5"]) ** 2)) > 1e-4
        f.reset()
        f.apply(dest, b)
        assert np.sqrt(np.mean(np.abs(dest - KAT["convolve_filter_first5"]) ** 2)) < 1e-4
        taps = oracle.complex_uniform(21, 777, dt)
        sig = oracle.complex_uniform(22, 20000, dt)
        g = kfr_amd.convolve_filter(taps, 256)
        out = np.empty_like(sig)
        pos = 0
        for chunk in (1, 255, 256, 1000, 3, 7000, 11485):
            g.apply(out[pos:pos + chunk], sig[pos:pos + chunk].copy())
            pos += chunk
        assert pos == sig.size
        ref = np.convolve(sig.astype(np.complex128), taps.astype(np.complex128))[: sig.size]
        tol = 1e-5 if dt == np.complex64 else 1e-12
        assert oracle.rel_rms(out, ref) < tol
        # new taps: the state restarts
        g.set_data(taps[:100])
        out2 = np.empty(500, dtype=dt)
        g.apply(out2, sig[:500].copy())
        ref2 = np.convolve(sig[:500].astype(np.complex128), taps[:100].astype(np.complex128))[:500]
        assert oracle.rel_rms(out2, ref2) < tol


def test_real_convolve_filter_set_data():
    taps = oracle.gaussian(31, 300, np.float32)
    sig = oracle.gaussian(32, 4000, np.float32)
    f = kfr_amd.convolve_filter(taps, 512)
    out = np.empty_like(sig)
    f.apply(out, sig)
    f.set_data(taps[:50])
    out2 = np.empty_like(sig)
    f.apply(out2, sig)
    ref2 = np.convolve(sig.astype(np.float64), taps[:50].astype(np.float64))[: sig.size]
    assert oracle.rel_rms(out2, ref2) < 1e-5
