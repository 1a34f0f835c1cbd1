"""Writes the golden fixtures of the parity tests from the reference engine.

Run in a container where oracle/_ref holds the reference build (`make -C oracle ref`):

    python tests/golden/make_golden.py

Outputs (data only: inputs and the reference's outputs):
  plans.json   reference plan descriptors (dump() fields, ISA suffix removed) for the config sizes
  perms.npz    reorder-stage permutations (iota through the reference's reorder stage)
  spectra.npz  seeded inputs and the reference's outputs for small transforms, real transforms,
               DCTs, convolutions and the streaming convolution filter
The known-answer vectors of the reference's own tests are in kat.json (transcribed values).
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

import oracle  # noqa: E402

# sizes that the configurations and the reference's tests use
PLAN_SIZES_C = sorted(set([1 << k for k in range(0, 25)] + list(range(2, 130)) + [
    210, 1029, 3150, 8000, 9990, 9988, 9999, 96001, 211, 3163, 15360, 30240, 28672, 15360000, 65536 * 3,
    4095, 2 * 97, 101 * 3, 127 * 127, 6000, 7 * 7 * 7 * 9 * 10]))
PLAN_SIZES_R = [2, 4, 6, 16, 30, 64, 420, 2048, 4096, 1 << 16, 1 << 20, 1 << 24, 30720000, 2 * 101, 2 * 3163]
PLAN_SIZES_DCT = [1, 2, 3, 16, 17, 64, 1000, 65536]


def strip_arch(s: str) -> str:
    return re.sub(r"\((avx2|avx512|sse2|sse41|avx|gfx950)\)", "", s)


def main():
    ref = oracle.reference()
    if ref is None:
        raise SystemExit("oracle/_ref is not built: run `make -C oracle ref`")
    plans = {"reference": strip_arch(ref.version())}
    for dt, tn in ((0, "f32"), (1, "f64")):
        for n in PLAN_SIZES_C:
            if dt == 1 and n > (1 << 22):
                continue
            plans[f"dft:{tn}:{n}"] = strip_arch(ref.describe(0, dt, n))
        for n in [1 << 12, 1 << 16, 1 << 20]:
            plans[f"dft_progressive:{tn}:{n}"] = strip_arch(ref.describe(3, dt, n))
        for n in PLAN_SIZES_R:
            for fmt in (0, 1):
                plans[f"real:{tn}:{n}:{fmt}"] = strip_arch(ref.describe(1, dt, n, fmt))
        for n in PLAN_SIZES_DCT:
            plans[f"dct:{tn}:{n}"] = strip_arch(ref.describe(2, dt, n))
    (HERE / "plans.json").write_text(json.dumps(plans, indent=0, sort_keys=True))

    perms = {}
    for dt, tn in ((0, "f32"), (1, "f64")):
        for n in [1 << 12, 1 << 13, 1 << 15, 1 << 16, 210, 1029, 3150, 8000, 9990, 360, 120, 7 * 7 * 7 * 9 * 10]:
            p = ref.permutation(dt, n)
            if p is not None:
                perms[f"{tn}_{n}"] = p.astype(np.int32)
    np.savez_compressed(HERE / "perms.npz", **perms)

    sp = {}
    for dt, cdt in ((np.float32, np.complex64), (np.float64, np.complex128)):
        tn = "f32" if dt == np.float32 else "f64"
        for n in [1, 2, 3, 5, 8, 16, 48, 60, 64, 97, 101, 127, 210, 1024, 4096]:
            x = oracle.complex_uniform(100 + n, n, cdt)
            sp[f"c_in_{tn}_{n}"] = x
            sp[f"c_fwd_{tn}_{n}"] = ref.dft(x, False)
            sp[f"c_inv_{tn}_{n}"] = ref.dft(x, True)
        # two of the 256 transforms of the 4096-point batch workload
        x = oracle.complex_uniform(4096, (2, 4096), cdt)
        sp[f"cfg1_in_{tn}"] = x
        sp[f"cfg1_fwd_{tn}"] = ref.dft(x, False)
        for n in [2, 4, 30, 64, 2048]:
            x = oracle.uniform(300 + n, n, dt)
            sp[f"r_in_{tn}_{n}"] = x
            for fmt in (0, 1):
                X = ref.rdft(x, fmt)
                sp[f"r_fwd_{tn}_{n}_{fmt}"] = X
                sp[f"r_inv_{tn}_{n}_{fmt}"] = ref.irdft(X, n, fmt)
        for n in [16, 17, 64]:
            x = oracle.uniform(500 + n, n, dt)
            sp[f"dct_in_{tn}_{n}"] = x
            sp[f"dct_fwd_{tn}_{n}"] = ref.dct(x, False)
            sp[f"dct_inv_{tn}_{n}"] = ref.dct(x, True)
        a = oracle.uniform(7, 100, dt)
        b = oracle.uniform(8, 37, dt)
        sp[f"conv_a_{tn}"] = a
        sp[f"conv_b_{tn}"] = b
        sp[f"conv_{tn}"] = ref.convolve(a, b, 0)
        sp[f"corr_{tn}"] = ref.convolve(a, b, 1)
        sp[f"acorr_{tn}"] = ref.convolve(a, None, 2)
        taps = oracle.uniform(9, 300, dt)
        sig = oracle.uniform(10, 5000, dt)
        f = ref.filter(taps, 64)
        chunks = [sig[:1], sig[1:700], sig[700:701], sig[701:3000], sig[3000:]]
        sp[f"filt_taps_{tn}"] = taps
        sp[f"filt_in_{tn}"] = sig
        sp[f"filt_out_{tn}"] = np.concatenate([f.apply(c) for c in chunks])
    ca = (oracle.uniform(11, 10, np.float32) + 1j * oracle.uniform(12, 10, np.float32)).astype(np.complex64)
    cb = (oracle.uniform(13, 6, np.float32) + 1j * oracle.uniform(14, 6, np.float32)).astype(np.complex64)
    sp["cconv_a"] = ca
    sp["cconv_b"] = cb
    sp["cconv"] = ref.convolve(ca, cb, 0)
    np.savez_compressed(HERE / "spectra.npz", **sp)
    print("wrote", ", ".join(p.name for p in (HERE / "plans.json", HERE / "perms.npz", HERE / "spectra.npz")))


if __name__ == "__main__":
    main()
