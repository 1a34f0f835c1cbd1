"""Builds libkfr_amd.so (the gfx950 DFT engine and its C ABI) in place.

    python -m kfr_amd.build            # compile for gfx950
    python -m kfr_amd.build --check    # print whether the library is up to date

All HIP code is compiled with packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
SRC = PKG / "csrc"
OBJ = PKG.parent / "_build"
LIB = PKG / "libkfr_amd.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++20", "-fPIC", "-fvisibility=hidden", "-Wall", "-Wno-unused-function",
          "-Wno-unknown-pragmas"]
HIPFLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-munsafe-fp-atomics"]
SOURCES = ["plan_desc.cpp", "capi.cpp", "fft_engine.hip", "fft_pow2.hip", "fft_batch.hip", "fft_md.hip", "fft_mixed.hip",
           "fft_noalloc.hip"]
HEADERS = ["plan_desc.hpp", "fft_engine.hpp", "fft_core.hpp", "fft_device.hpp", "fft_pow2.hpp", "fft_batch.hpp", "fft_md.hpp", "fft_mixed.hpp", "fft_timing.hpp"]


def _compile(src: str) -> Path:
    out = OBJ / (src.replace(".", "_") + ".o")
    cmd = [HIPCC] + COMMON + HIPFLAGS + ["-c", str(SRC / src), "-o", str(out)]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c",
                                  str(SRC / src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"compilation of {src} failed")
    return out


def up_to_date() -> bool:
    if not LIB.exists():
        return False
    t = LIB.stat().st_mtime
    deps = [SRC / s for s in SOURCES + HEADERS] + [PKG.parent / "include" / "kfr_amd_capi.h"]
    return all(d.stat().st_mtime <= t for d in deps)


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp)] + [str(o) for o in objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libkfr_amd.so failed")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--check", action="store_true")
    a = ap.parse_args()
    if a.check:
        print("up to date" if up_to_date() else "stale")
        return 0
    build()
    return 0


if __name__ == "__main__":
    sys.exit(main())
