"""Benchmark of the gfx950 DFT engine on the plain FP32 sweep workload.

Workload (one step): complex-f32 and complex-f64 C2C forward transforms at N = 2^4..2^24 with each
batch sized to 1 GiB of input (f32: 2^27 points, f64: 2^26 points), R2C dft_plan_real<float> at
N = 2^24 (batch 16) and N = 30,720,000 (batch 8), and DCT-II dct_plan<float> at N = 65536
(batch 1024). Inputs are i.i.d. standard Gaussian from a fixed seed, resident in HBM before timing.

value = whole-step GFLOP/s with 5 N log2 N per C2C transform and 2.5 N log2 N per R2C / DCT
transform. The line also carries the HBM roofline of the dominant kernel, the rel-RMS error of
sampled outputs against FP64, and the reference engine timed on the host CPU on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--quick]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure
ULP_EDGES = (0.5, 1.0, 2.0, 4.0, 8.0, 16.0, 64.0)  # ulp histogram bucket edges
# HBM bytes per dispatch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this benchmark
TRAFFIC_TABLE = "bench_sweep_hbm_traffic.md"


def sweep_configs(quick: bool):
    cfg = []
    top = 20 if quick else 24
    for a in range(4, top + 1):
        n = 1 << a
        cfg.append(("c2c", "f32", n, (1 << 27) // n))
    for a in range(4, top + 1):
        n = 1 << a
        cfg.append(("c2c", "f64", n, (1 << 26) // n))
    if not quick:
        cfg.append(("r2c", "f32", 1 << 24, 16))
        cfg.append(("r2c", "f32", 30720000, 8))
        cfg.append(("dct", "f32", 65536, 1024))
    return cfg


def flops(kind, n, batch):
    f = 5.0 * n * math.log2(n) * batch
    return f if kind == "c2c" else f / 2.0


def in_bytes_of(c):
    kind, dt, n, batch = c
    es = 4 if dt == "f32" else 8
    return (2 if kind == "c2c" else 1) * es * n * batch


def out_size_bytes(c):
    kind, dt, n, batch = c
    es = 4 if dt == "f32" else 8
    if kind == "c2c":
        return 2 * es * n * batch
    if kind == "r2c":
        return 2 * es * (n // 2 + 1) * batch
    return es * n * batch


def alg_bytes(kind, dt, n, batch):
    """Algorithmic HBM bytes: one read of the input and one write of the output."""
    es = 4 if dt == "f32" else 8
    if kind == "c2c":
        return 2 * (2 * es) * n * batch
    if kind == "r2c":
        return (es * n + 2 * es * (n // 2 + 1)) * batch
    return 2 * es * n * batch


def uneven_lengths() -> np.ndarray:
    """Every 7-smooth N in [18, 8192] and every prime in [17, 127] (328 lengths)."""
    smooth = {2 ** a * 3 ** b * 5 ** c * 7 ** d for a in range(14) for b in range(9) for c in range(6)
              for d in range(5)}
    smooth = {v for v in smooth if 18 <= v <= 8192}
    primes = [p for p in range(17, 128) if all(p % q for q in range(2, int(p ** 0.5) + 1))]
    return np.array(sorted(smooth | set(primes)), dtype=np.int64)


def uneven_batch(torch, kfr_amd, stream, reps: int, check: bool):
    """20,000 complex-f32 transforms with lengths drawn uniformly from uneven_lengths(), each with its
    own direction (Bernoulli 1/2), data U[-1,1) complex from a seed: one autotuned persistent graph
    launch against one plain-path launch per distinct (length, direction) on size-grouped data."""
    rng = np.random.default_rng(3)
    L = uneven_lengths()
    sizes = rng.choice(L, 20000)
    inv = rng.integers(0, 2, sizes.size).astype(np.int32)
    total = int(sizes.sum())
    x = (rng.uniform(-1, 1, 2 * total).astype(np.float32)).view(np.complex64)
    xd = torch.from_numpy(x.view(np.float32)).cuda()
    yd = torch.empty_like(xd)
    t0 = time.perf_counter()
    b = kfr_amd.dft_batch(sizes, inv, np.complex64, autotune=True)
    plan_s = time.perf_counter() - t0
    # per-size path: transforms grouped by (length, direction), contiguous per group
    groups = {}
    for i, (n, d) in enumerate(zip(sizes, inv)):
        groups.setdefault((int(n), int(d)), []).append(i)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    order = [i for k in sorted(groups) for i in groups[k]]
    xg = np.concatenate([x[offs[i]:offs[i] + sizes[i]] for i in order])
    xgd = torch.from_numpy(xg.view(np.float32)).cuda()
    ygd = torch.empty_like(xgd)
    plans = {n: kfr_amd.dft_plan(n, np.complex64) for (n, _) in groups}
    runs = []
    pos = 0
    for (n, d) in sorted(groups):
        cnt = len(groups[(n, d)])
        runs.append((plans[n], pos, n, cnt, bool(d)))
        pos += n * cnt

    def per_size():
        for p, o, n, cnt, d in runs:
            p.execute_batch(ygd[2 * o:], xgd[2 * o:], cnt, d, stream)

    def persistent():
        b.execute(yd, xd, stream)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    t_p = timed(persistent)
    t_s = timed(per_size)
    res = {"workload": "20000 complex-f32 transforms, lengths uniform over the 328 7-smooth N in [18,8192] "
                       "and primes in [17,127], direction Bernoulli(1/2), data U[-1,1) complex (seed 3)",
           "points": total, "ms_persistent": round(t_p, 4), "ms_per_size": round(t_s, 4),
           "launches_per_size": len(runs), "speedup": round(t_s / t_p, 2),
           "gbs_persistent": round(16.0 * total / (t_p * 1e-3) * 1e-9, 1),
           "frac_persistent": round(16.0 * total / (t_p * 1e-3) * 1e-9 / HBM_PEAK_GBS, 4),
           "plan_s": round(plan_s, 3), "schedule": b.describe()}
    if check:
        y = yd.cpu().numpy().view(np.complex64)
        worst = 0.0
        for i in rng.choice(sizes.size, 200, replace=False):
            n, o = int(sizes[i]), int(offs[i])
            xx = x[o:o + n].astype(np.complex128)
            ref = np.fft.ifft(xx) * n if inv[i] else np.fft.fft(xx)
            worst = max(worst, float(np.linalg.norm(y[o:o + n] - ref) / np.linalg.norm(ref)))
        res["rel_rms_vs_fp64_max_sampled"] = float(f"{worst:.3e}")
    return res


def precision_config(torch, kfr_amd, stream, reps: int):
    """Complex C2C N = 2^20, batch 64, unit-variance complex Gaussian data (seed 4) stored in bfloat16 and
    binary16 (radix-16 passes on MFMA) against the FP32 path on the same values; errors against the
    FP64 transform of the stored input, on 4 sampled transforms."""
    n, batch = 1 << 20, 64
    g = torch.Generator(device="cuda")
    g.manual_seed(4)
    x32 = torch.randn(batch, n, 2, device="cuda", generator=g) * (0.5 ** 0.5)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def errors(x, y):
        xc = torch.view_as_complex(x[:4].double().contiguous()).cpu().numpy()
        yc = torch.view_as_complex(y[:4].double().contiguous()).cpu().numpy()
        ref = np.fft.fft(xc, axis=-1)
        rel = np.sqrt(np.sum(np.abs(yc - ref) ** 2, axis=-1) / np.sum(np.abs(ref) ** 2, axis=-1))
        return float(f"{rel.max():.3e}"), float(f"{np.max(np.abs(yc - ref)) / np.max(np.abs(ref)):.3e}")

    res = {"workload": "complex C2C N=2^20 x 64, unit-variance complex Gaussian (seed 4), stored bf16 / f16 vs f32"}
    p32 = kfr_amd.dft_plan(n, np.complex64)
    out = {}
    for name, dt in (("f32", torch.float32), ("bf16", torch.bfloat16), ("f16", torch.float16)):
        x = x32.to(dt)
        y = torch.empty_like(x)
        if name == "f32":
            fn = lambda x=x, y=y: p32.execute_batch(y, x, batch, False, stream)  # noqa: E731
        else:
            ph = kfr_amd.dft_plan_half(n, name)
            fn = lambda x=x, y=y, ph=ph: ph.execute_batch(y, x, batch, False, stream)  # noqa: E731
        t = timed(fn)
        rel, mx = errors(x, y)
        bytes_alg = 2 * (8 if name == "f32" else 4) * n * batch
        out[name] = {"ms": round(t, 4), "gbs": round(bytes_alg / (t * 1e-3) * 1e-9, 1),
                     "frac": round(bytes_alg / (t * 1e-3) * 1e-9 / HBM_PEAK_GBS, 4),
                     "gflops": round(5.0 * n * 20 * batch / (t * 1e-3) * 1e-9, 1),
                     "rel_rms_vs_fp64_max": rel, "max_err_rel": mx}
        if name != "f32":
            out[name]["schedule"] = ph.describe()
    for name in ("bf16", "f16"):
        out[name]["speedup_vs_f32"] = round(out["f32"]["ms"] / out[name]["ms"], 3)
    res.update(out)
    # MFMA utilisation of the half-precision kernels of this workload, from the committed counter run
    # (rocprofv3 cannot run inside this process); per dispatch, in percent of the chip's SIMD cycles
    from kfr_amd.rocprof_summary import read_table
    mfile = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "half_c2c_2p20x64_mfma_counters.md")
    rows = [r for r in read_table(mfile) if r["kernel"].startswith("k_pow2<float") and r["kernel"].endswith((", 1>", ", 2>"))]
    if rows:
        res["mfma_counters"] = {"source": "profiles/half_c2c_2p20x64_mfma_counters.md",
                                "kernels": {r["kernel"]: {"mfma_util_pct": float(r["MfmaUtil"]),
                                                          "valu_busy_pct": float(r["VALUBusy"]),
                                                          "wait_share": float(r["wait share"])} for r in rows}}
    return res


def _timed(torch, stream, fn, reps):
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def round_trip_config(torch, kfr_amd, stream, reps: int):
    """Complex-f32 N = 4096, batch 256, re and im i.i.d. U[-1, 1) (seed 1): forward then inverse. The
    8 MiB batch stays in the 256 MB Infinity Cache, so this is a latency and throughput figure, not an
    HBM one."""
    n, batch = 4096, 256
    g = torch.Generator(device="cuda")
    g.manual_seed(1)
    x = torch.rand(batch * n * 2, device="cuda", generator=g) * 2 - 1
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    p = kfr_amd.dft_plan(n, np.complex64)

    def fn():
        p.execute_batch(y, x, batch, False, stream)
        p.execute_batch(z, y, batch, True, stream)

    t = _timed(torch, stream, fn, max(reps, 20))
    err = float((z / n - x).norm() / x.norm())
    return {"workload": "complex-f32 N=4096 x 256, U[-1,1) (seed 1), forward then inverse",
            "ms_round_trip": round(t, 4), "us_per_transform": round(t * 1e3 / (2 * batch), 4),
            "gflops": round(2 * flops("c2c", n, batch) / (t * 1e-3) * 1e-9, 1),
            "round_trip_rel_rms": float(f"{err:.3e}")}


def md_config(torch, kfr_amd, stream, reps: int):
    """Multidimensional complex-f32 transforms of 2^24 points: 4096 x 4096 and 256 x 256 x 256, forward,
    Gaussian data (seed 7). Algorithmic bytes: one read and one write of the array (16 bytes per point)."""
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    out = {}
    for shape in ((4096, 4096), (256, 256, 256)):
        n = int(np.prod(shape))
        x = torch.randn(2 * n, device="cuda", generator=g)
        y = torch.empty_like(x)
        p = kfr_amd.dft_plan_md(shape, np.complex64)
        t = _timed(torch, stream, lambda: p.execute_batch(y, x, 1, False, stream), reps)
        xs = x[: 2 * n].view(torch.complex64).reshape(shape)
        ref = torch.fft.fftn(xs.to(torch.complex128))
        err = float((y.view(torch.complex64).reshape(shape).to(torch.complex128) - ref).norm() / ref.norm())
        key = "x".join(str(v) for v in shape)
        out[key] = {"ms": round(t, 4), "gbs": round(16 * n / (t * 1e-3) * 1e-9, 1),
                    "frac": round(16 * n / (t * 1e-3) * 1e-9 / HBM_PEAK_GBS, 4),
                    "gflops": round(5.0 * n * math.log2(n) / (t * 1e-3) * 1e-9, 1),
                    "rel_rms_vs_fp64": float(f"{err:.3e}")}
    return out


def conv_config(torch, kfr_amd, stream, reps: int):
    """FFT convolution at full size: x = 128 signals of 2^20 f32 Gaussian samples (seed 5), h = 128
    Hann-windowed sincs of 4095 taps (cutoff 0.2 fs) plus 1e-3 N(0,1) (seed 6); y = x * h (length
    1,052,670), then the adjoint (dx, dh) of the same call, both by overlap-save in blocks of 2^14
    points (the filter spectra are recomputed in every call). Algorithmic bytes:
    forward 4 (Nx + Nh + Nx+Nh-1), adjoint 4 (Nx+Nh-1 + Nx + Nh + Nx + Nh) per signal. Errors against
    torch.fft (rocFFT) in float64 on 2 signals."""
    B, nx, nh = 128, 1 << 20, 4095
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    x = torch.randn(B, nx, device="cuda", generator=g)
    g.manual_seed(6)
    m = torch.arange(nh, device="cuda", dtype=torch.float64) - (nh - 1) / 2
    base = 0.4 * torch.sinc(0.4 * m) * torch.hann_window(nh, periodic=False, device="cuda", dtype=torch.float64)
    h = (base.float()[None, :] + 1e-3 * torch.randn(B, nh, device="cuda", generator=g)).contiguous()
    ny = nx + nh - 1
    gout = torch.randn(B, ny, device="cuda", generator=g)
    L = kfr_amd.lib()
    from kfr_amd.autograd import _plan
    plan = _plan(torch.float32, nx, nh)
    y = torch.empty(B, ny, device="cuda")
    dx = torch.empty_like(x)
    dh = torch.empty_like(h)
    st = stream.cuda_stream

    def fwd():
        assert L.kfr_amd_conv_forward(plan, y.data_ptr(), x.data_ptr(), h.data_ptr(), nh, B, 0, st) == 0

    def bwd():
        assert L.kfr_amd_conv_backward(plan, dx.data_ptr(), dh.data_ptr(), gout.data_ptr(), x.data_ptr(),
                                       h.data_ptr(), nh, B, 0, st) == 0

    tf = _timed(torch, stream, fwd, reps)
    tb = _timed(torch, stream, bwd, reps)
    bf = 4 * (nx + nh + ny) * B
    bb = 4 * (ny + nx + nh + nx + nh) * B
    # float64 reference on 2 signals
    S = 1 << 21
    xr = x[:2].double().requires_grad_(True)
    hr = h[:2].double().requires_grad_(True)
    yr = torch.fft.irfft(torch.fft.rfft(xr, S) * torch.fft.rfft(hr, S), S)[:, :ny]
    yr.backward(gout[:2].double())

    def rel(a, b):
        return float(f"{float((a.double() - b).norm() / b.norm()):.3e}")

    return {"workload": "128 x 2^20 f32 signals, 128 x 4095-tap Hann sinc (0.2 fs) + 1e-3 noise, "
                        "linear convolution and its adjoint (overlap-save, 2^14-point blocks)",
            "forward_ms": round(tf, 3), "forward_gbs": round(bf / (tf * 1e-3) * 1e-9, 1),
            "forward_frac": round(bf / (tf * 1e-3) * 1e-9 / HBM_PEAK_GBS, 4),
            "adjoint_ms": round(tb, 3), "adjoint_gbs": round(bb / (tb * 1e-3) * 1e-9, 1),
            "adjoint_frac": round(bb / (tb * 1e-3) * 1e-9 / HBM_PEAK_GBS, 4),
            "rel_rms_vs_fp64": {"y": rel(y[:2], yr.detach()), "dx": rel(dx[:2], xr.grad),
                                "dh": rel(dh[:2], hr.grad)}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--quick", action="store_true", help="sizes up to 2^20 only, C2C only")
    ap.add_argument("--only", default=None,
                    help="comma-separated subset of configurations kind:dtype:n (e.g. c2c:f32:4096), for profiling; "
                         "the words uneven, precision and extra add those measurements to the subset")
    ap.add_argument("--no-uneven", action="store_true", help="skip the mixed-length batch measurement")
    ap.add_argument("--no-precision", action="store_true", help="skip the half-precision measurement")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the round-trip (N=4096 x 256), full-size convolution and device-copy measurements")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run one plain 256 MiB device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU sharding is a design note only",
                          "n_gpus": args.gpus}))
        return 2

    import torch

    import kfr_amd
    from kfr_amd.rocprof_summary import read_traffic

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    cfgs = sweep_configs(args.quick)
    extras = set()
    if args.only:
        words = args.only.split(",")
        extras = {w for w in words if w in ("uneven", "precision", "extra")}
        keep = {tuple(x.split(":")) for x in words if x not in extras}
        cfgs = [c for c in cfgs if (c[0], c[1], str(c[2])) in keep]
        if not cfgs:
            print(json.dumps({"error": f"no configuration matches --only {args.only}"}))
            return 2

    # plans (excluded from timing)
    plans = {}
    for kind, dt, n, batch in cfgs:
        key = (kind, dt, n)
        if kind == "c2c":
            plans[key] = kfr_amd.dft_plan(n, np.complex64 if dt == "f32" else np.complex128)
        elif kind == "r2c":
            plans[key] = kfr_amd.dft_plan_real(n, kfr_amd.CCs, np.float32)
        else:
            plans[key] = kfr_amd.dct_plan(n, np.float32)

    # inputs resident in HBM: 1 GiB of Gaussian values, reinterpreted per configuration
    g = torch.Generator(device=dev)
    g.manual_seed(20240611)
    in_bytes = 1 << 30
    src32 = torch.randn(in_bytes // 4, generator=g, device=dev, dtype=torch.float32)
    src64 = torch.randn(in_bytes // 8, generator=g, device=dev, dtype=torch.float64)
    # the output buffer holds the largest output of any configuration (R2C writes n/2+1 complex values
    # per item, more than its input)
    out_bytes = max(out_size_bytes(c) for c in cfgs)
    for c in cfgs:
        assert in_bytes_of(c) <= in_bytes, c
    dst = torch.empty((out_bytes + 3) // 4, device=dev, dtype=torch.float32)

    if args.copy_probe:
        probe_src = src32[: (256 << 20) // 4]
        probe_dst = torch.empty_like(probe_src)
        probe_dst.copy_(probe_src)
        torch.cuda.synchronize()
        del probe_dst

    def run_cfg(c):
        kind, dt, n, batch = c
        p = plans[(kind, dt, n)]
        src = src32 if dt == "f32" else src64
        out = dst if dt == "f32" else dst.view(torch.float64)
        if kind == "c2c":
            p.execute_batch(out, src, batch, False, stream)
        elif kind == "r2c":
            p.execute_batch(out, src, batch, False, stream)
        else:
            p.execute_batch(out, src, batch, False, stream)

    for _ in range(args.warmup):
        for c in cfgs:
            run_cfg(c)
    torch.cuda.synchronize()

    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in cfgs]
          for _ in range(args.steps)]
    t0 = time.perf_counter()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    # every kernel launch of the timed region is also bracketed by HIP events on the same stream
    # (recorded by the library), giving each kernel's average launch duration
    with kfr_amd.kernel_timing() as kt:
        torch.cuda.synchronize()
        start.record(stream)
        for s in range(args.steps):
            for i, c in enumerate(cfgs):
                ev[s][i][0].record(stream)
                run_cfg(c)
                ev[s][i][1].record(stream)
        end.record(stream)
        torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_ms = start.elapsed_time(end)
    ms_per_step = total_ms / args.steps

    per = []
    for i, c in enumerate(cfgs):
        ts = [ev[s][i][0].elapsed_time(ev[s][i][1]) for s in range(args.steps)]
        t = float(np.median(ts)) * 1e-3
        kind, dt, n, batch = c
        per.append({"kind": kind, "dtype": dt, "n": n, "batch": batch, "ms": t * 1e3,
                    "gflops": flops(kind, n, batch) / t * 1e-9,
                    "gbs": alg_bytes(kind, dt, n, batch) / t * 1e-9,
                    "frac": alg_bytes(kind, dt, n, batch) / t * 1e-9 / HBM_PEAK_GBS})
    step_flops = sum(flops(c[0], c[2], c[3]) for c in cfgs)
    value = step_flops * args.steps / (total_ms * 1e-3) * 1e-9

    # rel-RMS against FP64 on sampled transforms
    errs = {}
    for c in cfgs:
        kind, dt, n, batch = c
        nsamp = max(1, min(batch, (1 << 18) // n))
        run_cfg(c)
        torch.cuda.synchronize()
        if kind == "c2c":
            cdt = np.complex64 if dt == "f32" else np.complex128
            src = (src32 if dt == "f32" else src64)
            x = src[: 2 * n * nsamp].cpu().numpy().view(cdt).reshape(nsamp, n)
            out = dst if dt == "f32" else dst.view(torch.float64)
            y = out[: 2 * n * nsamp].cpu().numpy().view(cdt).reshape(nsamp, n)
            ref = np.fft.fft(x.astype(np.complex128), axis=-1)
        elif kind == "r2c":
            x = src32[: n * nsamp].cpu().numpy().reshape(nsamp, n)
            cs = n // 2 + 1
            y = dst[: 2 * cs * nsamp].cpu().numpy().view(np.complex64).reshape(nsamp, cs)
            ref = np.fft.rfft(x.astype(np.float64), axis=-1)
        else:
            import scipy.fft
            x = src32[: n * nsamp].cpu().numpy().reshape(nsamp, n)
            y = dst[: n * nsamp].cpu().numpy().reshape(nsamp, n)
            ref = scipy.fft.dct(x.astype(np.float64), type=2, axis=-1) / 2.0
        rel = [float(np.linalg.norm(y[b] - ref[b]) / np.linalg.norm(ref[b])) for b in range(nsamp)]
        k = f"{kind}_{dt}"
        a = errs.setdefault(k, {"max": 0.0, "sum": 0.0, "count": 0, "max_abs_rel": 0.0})
        a["max"] = max(a["max"], max(rel))
        a["sum"] += sum(rel)
        a["count"] += len(rel)
        a["max_abs_rel"] = max(a["max_abs_rel"], float(np.max(np.abs(y - ref)) / np.max(np.abs(ref))))
        if k == "c2c_f32":
            # error of each real component in units in the last place of its FP64 value rounded to f32
            yr = np.stack([y.real, y.imag]).ravel().astype(np.float64)
            rr = np.stack([ref.real, ref.imag]).ravel()
            ulp = np.spacing(np.abs(rr).astype(np.float32)).astype(np.float64)
            u = np.abs(yr - rr) / ulp
            hist = errs.setdefault("ulp_hist", np.zeros(len(ULP_EDGES) + 1, dtype=np.int64))
            hist += np.bincount(np.searchsorted(ULP_EDGES, u, side="right"), minlength=len(ULP_EDGES) + 1)
    ulp_hist = errs.pop("ulp_hist", None)
    err_report = {k: {"rel_rms_max": float(f"{v['max']:.3e}"), "rel_rms_mean": float(f"{v['sum'] / v['count']:.3e}"),
                      "max_abs_rel": float(f"{v['max_abs_rel']:.3e}"), "transforms": v["count"]}
                  for k, v in errs.items()}
    if ulp_hist is not None:
        labels = ([f"<{ULP_EDGES[0]:g}"] + [f"{ULP_EDGES[i]:g}-{ULP_EDGES[i + 1]:g}" for i in range(len(ULP_EDGES) - 1)]
                  + [f">={ULP_EDGES[-1]:g}"])
        err_report["ulp_histogram_c2c_f32"] = {lab: int(c) for lab, c in zip(labels, ulp_hist)}

    # a plain device copy of 1 GiB: the bandwidth a streaming kernel reaches on this device (not in
    # profiling runs, whose calibration probe is the largest copy)
    copy_gbs = None
    if not args.no_extra:
        copy_dst = torch.empty_like(src32)
        copy_ms = _timed(torch, stream, lambda: copy_dst.copy_(src32), 5)
        copy_gbs = 2 * src32.numel() * 4 / (copy_ms * 1e-3) * 1e-9
        del copy_dst

    # dominant kernel: the kernel (template instance) with the largest total time in the timed region;
    # achieved = its algorithmic bytes per launch / its average launch duration
    kernels = kt.by_kernel()
    top = []
    for name, k in list(kernels.items())[: (None if args.only else 8)]:
        avg_ms = k["ms"] / k["launches"]
        gbs = k["bytes"] / (k["ms"] * 1e-3) * 1e-9 if k["ms"] > 0 else 0.0
        top.append({"kernel": name, "launches_per_step": k["launches"] / args.steps, "avg_us": round(avg_ms * 1e3, 1),
                    "alg_bytes_per_launch": k["bytes"] // k["launches"], "gbs": round(gbs, 1),
                    "frac": round(gbs / HBM_PEAK_GBS, 4), "share_of_step": round(k["ms"] / total_ms, 4)})
    roofline = None
    if top:
        d = top[0]
        tfile = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", TRAFFIC_TABLE)
        tr = read_traffic(tfile, d["kernel"])
        roofline = {"bound": "hbm", "achieved": d["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": d["frac"],
                    "traffic": int(tr) if tr is not None else None,
                    "kernel": d["kernel"], "avg_us": d["avg_us"], "alg_bytes_per_launch": d["alg_bytes_per_launch"],
                    "launches_per_step": d["launches_per_step"], "share_of_step": d["share_of_step"],
                    "traffic_source": f"profiles/{TRAFFIC_TABLE}" if tr is not None else None,
                    "copy_gbs": round(copy_gbs, 1) if copy_gbs else None,
                    "frac_of_copy": round(d["gbs"] / copy_gbs, 4) if copy_gbs else None}
    # f32 lengths up to 2^15 run in one workgroup per transform, longer ones in two or three HBM passes
    fs = [p for p in per if p["kind"] == "c2c" and p["dtype"] == "f32" and p["n"] > (1 << 15)]
    fs_t = sum(p["ms"] for p in fs)
    fs_bytes = sum(alg_bytes("c2c", "f32", p["n"], p["batch"]) for p in fs)
    lds = [p for p in per if p["kind"] == "c2c" and p["dtype"] == "f32" and p["n"] <= (1 << 15)]
    lds_t = sum(p["ms"] for p in lds)
    lds_bytes = sum(alg_bytes("c2c", "f32", p["n"], p["batch"]) for p in lds)

    uneven = None
    if not args.no_uneven and (not args.only or "uneven" in extras):
        uneven = uneven_batch(torch, kfr_amd, stream, max(3, args.steps), check=True)

    precision = None
    if not args.no_precision and (not args.only or "precision" in extras):
        precision = precision_config(torch, kfr_amd, stream, max(3, args.steps))

    round_trip = conv = md = None
    if not args.no_extra and (not args.only or "extra" in extras):
        round_trip = round_trip_config(torch, kfr_amd, stream, max(3, args.steps))
        conv = conv_config(torch, kfr_amd, stream, max(3, args.steps))
        md = md_config(torch, kfr_amd, stream, max(3, args.steps))

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(cfgs)

    line = {
        "metric": "batched C2C FFT GFLOP/s (5N*log2N) & % HBM roofline; rel-RMS err vs FP64",
        "value": round(value, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (i.i.d. standard Gaussian, torch generator seed 20240611, resident in HBM)",
        "config": {"workload": "plain sweep: C2C c32+c64 N=2^4..2^24 at 1 GiB input per size"
                               + ("" if args.quick else ", R2C f32 N=2^24 x16 and N=30720000 x8, DCT-II f32 N=65536 x1024"),
                   "parallelism": "single GPU"},
        # the long per-configuration lists first and the summaries last, so that a reader of the line's
        # tail (a log that keeps its last few KiB) still sees the roofline and the extra measurements
        "per_config": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in p.items()} for p in per],
        "top_kernels": top,
        "errors_vs_fp64": err_report,
        "wall_s": round(wall, 3),
        "single_kernel_c2c_f32_frac": round(lds_bytes / (lds_t * 1e-3) * 1e-9 / HBM_PEAK_GBS, 4) if lds_t else None,
        "long_c2c_f32_frac": round(fs_bytes / (fs_t * 1e-3) * 1e-9 / HBM_PEAK_GBS, 4) if fs_t else None,
        "md_c32": md,
        "round_trip_4096x256": round_trip,
        "conv": conv,
        "precision": precision,
        "uneven_batch": uneven,
        "cpu_baseline": cpu,
        "roofline": roofline,
    }
    print(json.dumps(line))
    return 0


def cpu_baseline(cfgs):
    """The reference engine (oracle/_ref, built from the reference sources) on the host cores, timed as
    the reference's benchmarking guidance asks: one shared plan per point, the batch split across the
    threads (each with its own temp buffer), planning excluded, at least 1 s per point as the median of
    5 repetitions of at least 0.2 s. The sample is 10 points of the sweep (complex f32 and f64 at short,
    medium and long lengths, R2C f32 at 2^24, DCT-II f32 at 65536); every configuration of the sweep is
    charged at the rate of the sampled point of its kind and precision nearest in log2 N, and the value
    is the sweep's flops over the time that gives, the weighting of the GPU figure. Threads: the box's
    CPU share (16 for one GPU) or the cores this process may use, whichever is smaller."""
    try:
        import oracle
    except ImportError:
        return None
    ref = oracle.reference()
    if ref is None:
        return {"value": None, "unit": "GFLOP/s", "cores": 0, "kind": "reference",
                "sample": "skipped: oracle/_ref not built"}
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    threads = max(1, min(16, avail))
    sample = [("c2c", "f32", 1 << 10), ("c2c", "f32", 1 << 14), ("c2c", "f32", 1 << 18), ("c2c", "f32", 1 << 22),
              ("c2c", "f32", 1 << 24), ("c2c", "f64", 1 << 10), ("c2c", "f64", 1 << 18), ("c2c", "f64", 1 << 24),
              ("r2c", "f32", 1 << 24), ("dct", "f32", 65536)]
    kind_code = {"c2c": 0, "r2c": 1, "dct": 2}
    rates = {}
    points = []
    # one pool of standard Gaussian values (1 GiB) and one output buffer serve every point
    pool = np.resize(oracle.gaussian(11, 1 << 22, np.float32), 1 << 28)
    outbuf = np.empty(1 << 28, np.float32)
    for kind, dt, n in sample:
        real = kind != "c2c"
        fdt = np.float32 if dt == "f32" else np.float64
        cdt = np.complex64 if dt == "f32" else np.complex128
        es = np.dtype(fdt if real else cdt).itemsize
        cap = max(1, (1 << 30) // (n * es))                   # at most 1 GiB of input

        def run(batch, reps):
            x = pool.view(fdt)[: batch * n * (1 if real else 2)].view(fdt if real else cdt).reshape(batch, n)
            on = batch * (n // 2 + 1) if kind == "r2c" else batch * n
            out = outbuf.view(cdt if kind != "dct" else fdt)[:on]
            return ref.time(kind_code[kind], 0 if dt == "f32" else 1, n, batch, False, min(threads, batch), reps,
                            out, x)

        b = min(cap, threads)
        t = run(b, 1)
        while t < 0.2 and b < cap:                             # grow the batch to >= 0.2 s per repetition
            b = min(cap, max(b + 1, int(b * 0.25 / max(t, 1e-6))))
            t = run(b, 1)
        t = run(b, 5)
        r = flops(kind, n, b) / t
        rates[(kind, dt, n)] = r
        points.append(f"{kind} {dt} N=2^{int(math.log2(n))} x{b}: {r * 1e-9:.2f}")

    def rate_for(kind, dt, n):
        cands = [(abs(math.log2(k[2]) - math.log2(n)), v) for k, v in rates.items() if k[0] == kind and k[1] == dt]
        if not cands:
            cands = [(abs(math.log2(k[2]) - math.log2(n)), v) for k, v in rates.items() if k[0] == "c2c" and k[1] == dt]
        return min(cands)[1]

    tot_f = sum(flops(k, n, b) for k, dt, n, b in cfgs)
    tot_t = sum(flops(k, n, b) / rate_for(k, dt, n) for k, dt, n, b in cfgs)
    return {"value": round(tot_f / tot_t * 1e-9, 2), "unit": "GFLOP/s", "cores": threads, "kind": "reference",
            "arch": ref.version(),
            "sample": "reference dft_plan / dft_plan_real / dct_plan on " + str(threads) + " threads, median of 5 "
                      "repetitions of >= 0.2 s per point, GFLOP/s per point: " + "; ".join(points) +
                      "; the sweep's configurations charged at the nearest sampled rate"}


if __name__ == "__main__":
    sys.exit(main())
