"""The engine's HIP kernels run on the CPU through an emulation of the HIP runtime (tests/cpu_emu):
every workgroup as host threads meeting at its barriers, its LDS as one heap block of exactly the
launch's size, device buffers as exact host allocations. Built with AddressSanitizer, any LDS or
global access outside its buffer aborts the run; the spectra are checked against numpy in FP64.

Cases: the generic four-step over factors 3, 5, 7 with workgroups that straddle batch items
(f32 at 16 points per thread), the power-of-two four-step, the mixed-radix menu lengths, Bluestein, a
real transform, a DCT, the one-kernel DCT-II/III and R2C/C2R, the R2C repack pass; half-precision storage on emulated MFMA; and the f64
generic kernels built at 16 points per thread (the geometry an earlier build ran on the GPU, where
lengths with radix-5 passes came out wrong) on those lengths.
Set KFR_AMD_EMU_TSAN=1 to also run a ThreadSanitizer build (a missing barrier shows as a race), and
KFR_AMD_EMU_FULL=1 for the 16-points-per-thread f64 build and the slowest f64 length."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE / "cpu_emu"))
import build as emu_build  # noqa: E402

import oracle  # noqa: E402
from tests.common import fp64_dct, fp64_rdft  # noqa: E402

pytestmark = pytest.mark.skipif(not Path(emu_build.CXX).exists(), reason="ROCm clang++ not found")


def _run(variant, tmp_path, args, x, out_dtype, env=None):
    exe = emu_build.build(variant)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    np.ascontiguousarray(x).tofile(fin)
    r = subprocess.run([str(exe)] + [str(a) for a in args] + [str(fin), str(fout)], capture_output=True, text=True,
                       timeout=900, env=None if env is None else dict(os.environ, **env))
    assert r.returncode == 0, r.stderr[-4000:]
    return np.fromfile(fout, dtype=out_dtype)


def _dft(variant, tmp_path, dtype, n, batch, inverse, seed):
    t = "f32" if dtype == np.complex64 else "f64"
    x = oracle.complex_uniform(seed, (batch, n), dtype)
    y = _run(variant, tmp_path, ["dft", t, n, batch, int(inverse)], x, dtype).reshape(batch, n)
    xx = x.astype(np.complex128)
    ref = np.fft.ifft(xx, axis=1) * n if inverse else np.fft.fft(xx, axis=1)
    return max(oracle.rel_rms(y[b], ref[b]) for b in range(batch))


# (n, batch): 59049 = 3^10 and 30240 = 2^5 3^3 5 7 run the generic four-step with odd column and row
# counts; 13440 = 2^7 3 5 7 and 62500 = 2^2 5^6 are the lengths that failed at 16 f64 points per
# thread; 65536 the power-of-two four-step; 3840 a menu length; 10007 chained Bluestein (M = 2^15); 97 a direct
# prime pass
F32_CASES = [(59049, 3), (30240, 3), (13440, 2), (1 << 16, 2), (3840, 3), (10007, 1), (97, 5)]
F64_CASES = [(13440, 3), (15360, 2), (30240, 3), (62500, 1)]
# the default run keeps the CPU suite to a few minutes: one sanitizer build, and f64 without its
# slowest length; KFR_AMD_EMU_FULL=1 adds 15360 and the 16-points-per-thread f64 build
FULL = os.environ.get("KFR_AMD_EMU_FULL") == "1"
F64_VARIANTS = ["asan", "asan_pm16"] if FULL else ["asan"]
if not FULL:
    F64_CASES = [c for c in F64_CASES if c[0] != 15360]


@pytest.mark.parametrize("n,batch", F32_CASES)
def test_emulated_f32_lengths(n, batch, tmp_path):
    for inverse in (False, True):
        e = _dft("asan", tmp_path, np.complex64, n, batch, inverse, 40 + n % 97)
        print(f"emulated f32 n={n} batch={batch} inv={inverse}: rel_rms_fp64={e:.3e}", flush=True)
        assert e <= 1e-6


@pytest.mark.parametrize("variant", F64_VARIANTS)
@pytest.mark.parametrize("n,batch", F64_CASES)
def test_emulated_f64_lengths(variant, n, batch, tmp_path):
    for inverse in (False, True):
        e = _dft(variant, tmp_path, np.complex128, n, batch, inverse, 50 + n % 89)
        print(f"emulated f64 ({variant}) n={n} batch={batch} inv={inverse}: rel_rms_fp64={e:.3e}", flush=True)
        assert e <= 1e-14


def test_emulated_real_and_dct(tmp_path):
    n, batch = 1 << 14, 3
    x = oracle.gaussian(71, batch * n, np.float32).reshape(batch, n)
    for fmt in (0, 1):
        X = _run("asan", tmp_path, ["real", "f32", n, fmt, batch, 0], x, np.complex64).reshape(batch, -1)
        assert oracle.rel_rms(X, fp64_rdft(x.astype(np.float64), fmt)) <= 1e-6
        y = _run("asan", tmp_path, ["real", "f32", n, fmt, batch, 1], X, np.float32).reshape(batch, n)
        assert oracle.rel_rms(y / n, x) <= 1e-6
    m = 4096
    d = oracle.gaussian(72, 2 * m, np.float32).reshape(2, m)
    for inverse in (False, True):
        y = _run("asan", tmp_path, ["dct", "f32", m, 2, int(inverse)], d, np.float32).reshape(2, m)
        assert oracle.rel_rms(y, fp64_dct(d.astype(np.float64), inverse)) <= 2e-6


# R2C and C2R in one kernel for power-of-two halves that fit one workgroup (f32 2^8, 2^10, 2^13 and
# 2^15 halves, the last two with two-round LDS exchanges; f64 2^14), both pack formats
@pytest.mark.parametrize("t,n,batch", [("f32", 512, 5), ("f32", 2048, 3), ("f32", 1 << 14, 2), ("f32", 1 << 16, 2),
                                       ("f64", 1 << 15, 2)])
def test_emulated_real_one_kernel(t, n, batch, tmp_path):
    dt = np.float32 if t == "f32" else np.float64
    cdt = np.complex64 if t == "f32" else np.complex128
    bound = 1e-6 if t == "f32" else 1e-14
    for fmt in (0, 1):
        x = oracle.gaussian(75 + fmt, batch * n, dt).reshape(batch, n)
        X = _run("asan", tmp_path, ["real", t, n, fmt, batch, 0], x, cdt).reshape(batch, -1)
        ref = fp64_rdft(x.astype(np.float64), fmt)
        assert oracle.rel_rms(X, ref) <= bound
        y = _run("asan", tmp_path, ["real", t, n, fmt, batch, 1], ref.astype(cdt), dt).reshape(batch, n)
        assert oracle.rel_rms(y / n, x) <= bound


# C2R of lengths whose half does not fit one generic workgroup: the unpacking pass into plain
# half-length sequences, then the power-of-two kernels (2^16) or the generic four-step (30000)
@pytest.mark.parametrize("n,batch,fmt", [(1 << 16, 2, 1), (1 << 16, 3, 0), (30000, 2, 1)])
def test_emulated_c2r_unpack_pass(n, batch, fmt, tmp_path):
    x = oracle.gaussian(74 + fmt, batch * n, np.float32).reshape(batch, n)
    X = fp64_rdft(x.astype(np.float64), fmt).astype(np.complex64)
    y = _run("asan", tmp_path, ["real", "f32", n, fmt, batch, 1], X, np.float32).reshape(batch, n)
    assert oracle.rel_rms(y / n, x) <= 1e-6


# R2C whose half does not fit one workgroup: the complex 2^16-point half transform, then the repack
# pass over the half spectrum in global memory
@pytest.mark.parametrize("fmt", [0, 1])
def test_emulated_r2c_repack_pass(fmt, tmp_path):
    n, batch = 1 << 17, 1
    x = oracle.gaussian(76 + fmt, batch * n, np.float32).reshape(batch, n)
    X = _run("asan", tmp_path, ["real", "f32", n, fmt, batch, 0], x, np.complex64).reshape(batch, -1)
    assert oracle.rel_rms(X, fp64_rdft(x.astype(np.float64), fmt)) <= 1e-6


# DCT-II and DCT-III in one kernel: f32 at 2^8 points per half (16 threads per transform, the shortest
# fused length), 2^10 (mirrored first pass, LDS partner exchange) and 2^13, 2^15 (two-round LDS
# exchange, mirrored first and last passes), f64 at 2^13 and 2^14; 128 points f32 stays on the
# pack/FFT/post path
@pytest.mark.parametrize("t,m,batch", [("f32", 512, 5), ("f32", 256, 3), ("f32", 2048, 3), ("f32", 1 << 14, 2),
                                       ("f32", 1 << 16, 1), ("f64", 1 << 14, 2), ("f64", 1 << 15, 1)])
def test_emulated_fused_dct(t, m, batch, tmp_path):
    dt = np.float32 if t == "f32" else np.float64
    d = oracle.gaussian(73 + m % 7, batch * m, dt).reshape(batch, m)
    for inverse in (False, True):
        y = _run("asan", tmp_path, ["dct", t, m, batch, int(inverse)], d, dt).reshape(batch, m)
        e = oracle.rel_rms(y, fp64_dct(d.astype(np.float64), inverse))
        print(f"emulated dct {t} m={m} batch={batch} inverse={inverse}: rel_rms_fp64={e:.3e}", flush=True)
        assert e <= (1e-6 if t == "f32" else 1e-14)


def _to_half(x: np.ndarray, storage: str) -> np.ndarray:
    """complex values as interleaved 2-byte pairs: binary16, or bfloat16 rounded to nearest even"""
    f = np.stack([x.real, x.imag], -1).astype(np.float32)
    if storage == "f16":
        return f.astype(np.float16).view(np.uint16)
    u = f.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def _from_half(h: np.ndarray, storage: str) -> np.ndarray:
    if storage == "f16":
        f = h.view(np.float16).astype(np.float64)
    else:
        f = (h.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    return f[..., 0] + 1j * f[..., 1]


# half-precision storage with the radix-16 passes on (emulated) MFMA: one workgroup (2^10), the
# four-step (2^15) and the multi-pass kernels with their scaled binary16 intermediates (2^15 and 2^16,
# selected through KFR_AMD_HALF_MULTIPASS_MIN_LOG)
@pytest.mark.parametrize("storage,n,batch,multi", [("bf16", 1 << 10, 3, False), ("bf16", 1 << 15, 2, False),
                                                   ("bf16", 1 << 15, 2, True), ("f16", 1 << 16, 1, True)])
def test_emulated_half_storage(storage, n, batch, multi, tmp_path):
    exe = emu_build.build("asan")
    rng = np.random.default_rng(n + batch)
    x = rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))
    h = _to_half(x, storage)
    xq = _from_half(h, storage)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    h.tofile(fin)
    env = dict(os.environ, KFR_AMD_HALF_MULTIPASS_MIN_LOG="15" if multi else "99")
    for inverse in (False, True):
        r = subprocess.run([str(exe), "half", storage, str(n), str(batch), str(int(inverse)), str(fin), str(fout)],
                           capture_output=True, text=True, timeout=900, env=env)
        assert r.returncode == 0, r.stderr[-4000:]
        y = _from_half(np.fromfile(fout, dtype=np.uint16).reshape(batch, n, 2), storage)
        ref = np.fft.ifft(xq, axis=1) * n if inverse else np.fft.fft(xq, axis=1)
        e = np.linalg.norm(y - ref) / np.linalg.norm(ref)
        print(f"emulated {storage} n={n} multi={multi} inv={inverse}: rel_rms_fp64={e:.3e}", flush=True)
        assert e <= (2e-3 if storage == "bf16" else 5e-4)


@pytest.mark.skipif(os.environ.get("KFR_AMD_EMU_TSAN") != "1", reason="set KFR_AMD_EMU_TSAN=1")
@pytest.mark.parametrize("dtype,n,batch", [(np.complex64, 30240, 3), (np.complex64, 1 << 15, 1),
                                           (np.complex128, 13440, 2)])
def test_emulated_no_races(dtype, n, batch, tmp_path):
    e = _dft("tsan", tmp_path, dtype, n, batch, False, 60)
    assert e <= (1e-6 if dtype == np.complex64 else 1e-14)
