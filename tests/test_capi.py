"""The C ABI boundary (CPU): the library loads, exports every symbol its header declares and the
reference's C API names, and its error channel behaves as the reference's (src/capi/capi.cpp:36-91)."""
from __future__ import annotations

import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

import kfr_amd
from kfr_amd import _lib

ROOT = Path(__file__).resolve().parent.parent

# the reference's C API entry points on the DFT path (include/kfr/capi.h:116-736)
REFERENCE_NAMES = [
    "kfr_version_string", "kfr_version", "kfr_enabled_archs", "kfr_current_arch", "kfr_last_error",
    "kfr_allocate", "kfr_allocate_aligned", "kfr_deallocate",
] + [f"kfr_dft_{f}_{t}" for t in ("f32", "f64") for f in (
    "create_plan", "create_2d_plan", "create_3d_plan", "create_md_plan", "dump", "get_size", "get_temp_size",
    "execute", "execute_inverse", "delete_plan")] + [f"kfr_dft_real_{f}_{t}" for t in ("f32", "f64") for f in (
        "create_plan", "create_2d_plan", "create_3d_plan", "create_md_plan", "dump", "get_size", "get_temp_size",
        "execute", "execute_inverse", "delete_plan")] + [f"kfr_dct_{f}_{t}" for t in ("f32", "f64") for f in (
            "create_plan", "dump", "get_size", "get_temp_size", "execute", "execute_inverse", "delete_plan")] + [
    f"kfr_filter_{f}_{t}" for t in ("f32", "f64") for f in (
        "create_fir_plan", "create_convolution_plan", "create_iir_plan", "process", "reset", "delete_plan")]


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_library_loads():
    L = kfr_amd.lib()
    assert L.kfr_version() == 70001
    assert b"gfx950" in L.kfr_version_string()
    assert L.kfr_current_arch() == 100


def test_every_declared_symbol_is_exported():
    decl = _lib.declared_symbols()
    assert len(decl) > 80
    missing = sorted(set(decl) - exported())
    assert not missing, missing


def test_reference_c_api_names_are_all_present():
    missing = sorted(set(REFERENCE_NAMES) - exported())
    assert not missing, missing
    decl = set(_lib.declared_symbols())
    assert set(REFERENCE_NAMES) <= decl


def test_header_cites_reference_interface():
    text = (ROOT / "include" / "kfr_amd_capi.h").read_text()
    assert "include/kfr/capi.h:207-372" in text and "include/kfr/capi.h:380-516" in text
    assert re.search(r"include/kfr/capi.h:523-636", text)


def test_header_compiles_as_c():
    src = '#include "kfr_amd_capi.h"\nint main(void){ return KFR_HEADERS_VERSION == 70000 ? 0 : 1; }\n'
    r = subprocess.run(["gcc", "-x", "c", "-std=c11", "-I", str(ROOT / "include"), "-fsyntax-only", "-"],
                       input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_allocation_alignment():
    L = kfr_amd.lib()
    p = L.kfr_allocate(1000)
    assert p and p % 64 == 0
    L.kfr_deallocate(p)
    q = L.kfr_allocate_aligned(100, 4096)
    assert q and q % 4096 == 0
    L.kfr_deallocate(q)


def test_error_channel_without_gpu():
    L = kfr_amd.lib()
    # odd real size: rejected before any device work (include/kfr/dft/fft.hpp:485)
    assert not L.kfr_dft_real_create_plan_f64(9, 1)
    assert "even" in kfr_amd.last_error()
    # NULL plan
    assert L.kfr_dft_get_size_f32(None) == 0
    assert "NULL" in kfr_amd.last_error()
    # FIR / IIR are not on the DFT path
    assert not L.kfr_filter_create_fir_plan_f32(None, 0)
    assert "FIR" in kfr_amd.last_error()
    assert not L.kfr_filter_create_iir_plan_f64(None, 0)
    assert "IIR" in kfr_amd.last_error()
    # success clears
    buf = C.create_string_buffer(256)
    assert L.kfr_amd_describe(0, 0, 8, 0, buf, 256) > 0
    assert kfr_amd.last_error() == ""
    # bad arguments
    assert L.kfr_amd_describe(9, 0, 8, 0, buf, 256) == -1
    assert "kind" in kfr_amd.last_error()


def test_error_text_is_per_thread():
    import threading
    L = kfr_amd.lib()
    assert not L.kfr_dft_real_create_plan_f32(3, 1)
    seen = []

    def other():
        seen.append(kfr_amd.last_error())

    t = threading.Thread(target=other)
    t.start()
    t.join()
    assert seen == [""]
    assert "even" in kfr_amd.last_error()


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "libkfr_amd.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError, match="no CPU fallback"):
        _lib.lib()
