"""Builds the engine's kernels and host code for the CPU emulation of HIP (hip/hip_runtime.h beside
this file) with a sanitizer, into _build/cpu_emu/<variant>/emu_main:

    python tests/cpu_emu/build.py asan        # AddressSanitizer, the library's geometry
    python tests/cpu_emu/build.py asan_pm16   # AddressSanitizer, f64 generic kernels at 16 points/thread
    python tests/cpu_emu/build.py tsan        # ThreadSanitizer

Host code only: ROCm's clang++ in x86 mode compiles the .hip sources as C++ against the emulation
header; nothing here targets the GPU.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
SRC = ROOT / "kfr_amd" / "csrc"
CXX = os.environ.get("KFR_AMD_EMU_CXX", "/opt/rocm/llvm/bin/clang++")
SOURCES = ["plan_desc.cpp", "fft_engine.hip", "fft_pow2.hip", "fft_mixed.hip"]
VARIANTS = {
    "asan": ["-fsanitize=address"],
    "asan_pm16": ["-fsanitize=address", "-DKFR_AMD_PMAX_F64=16"],
    "tsan": ["-fsanitize=thread"],
}


def exe_path(variant: str) -> Path:
    return ROOT / "_build" / "cpu_emu" / variant / "emu_main"


def _inputs() -> list[Path]:
    return [SRC / s for s in SOURCES] + list(SRC.glob("*.hpp")) + [HERE / "emu_main.cpp", HERE / "hip" / "hip_runtime.h",
                                                                  Path(__file__)]


def up_to_date(variant: str) -> bool:
    exe = exe_path(variant)
    if not exe.exists():
        return False
    t = exe.stat().st_mtime
    return all(p.stat().st_mtime <= t for p in _inputs())


def build(variant: str = "asan", jobs: int = 8) -> Path:
    flags = ["-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer", "-I", str(HERE)] + VARIANTS[variant]
    out = exe_path(variant).parent
    out.mkdir(parents=True, exist_ok=True)
    if up_to_date(variant):
        return exe_path(variant)

    def compile_one(src: Path) -> Path:
        obj = out / (src.name.replace(".", "_") + ".o")
        cmd = [CXX] + flags + ["-x", "c++", "-c", str(src), "-o", str(obj)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError(f"emulation build of {src.name} failed:\n{r.stderr[-4000:]}")
        return obj

    srcs = [SRC / s for s in SOURCES] + [HERE / "emu_main.cpp"]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    exe = exe_path(variant)
    r = subprocess.run([CXX] + VARIANTS[variant] + ["-o", str(exe)] + [str(o) for o in objs] + ["-lpthread"],
                       capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"emulation link failed:\n{r.stderr[-4000:]}")
    return exe


if __name__ == "__main__":
    for v in sys.argv[1:] or ["asan"]:
        print(build(v))
