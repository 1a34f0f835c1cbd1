"""Host-side logic of bench.py and the profile summaries it reads (cpu): the multi-GPU refusal, the
algorithmic-byte and flop counts, and the kernel-name and traffic-table parsing of the roofline."""
from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402
from kfr_amd import rocprof_summary as rs  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       cwd=ROOT, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_algorithmic_bytes_and_flops():
    n, b = 1 << 12, 256
    assert bench.alg_bytes("c2c", "f32", n, b) == 16 * n * b
    assert bench.alg_bytes("c2c", "f64", n, b) == 32 * n * b
    assert bench.alg_bytes("r2c", "f32", n, b) == (4 * n + 8 * (n // 2 + 1)) * b
    assert bench.alg_bytes("dct", "f32", n, b) == 8 * n * b
    assert bench.flops("c2c", n, b) == pytest.approx(5 * n * 12 * b)
    assert bench.flops("r2c", n, b) == pytest.approx(2.5 * n * 12 * b)
    # every sweep configuration fits the 1 GiB input buffer
    for c in bench.sweep_configs(False):
        assert bench.in_bytes_of(c) <= 1 << 30, c


def test_kernel_names_match_the_timing_records():
    assert rs.kernel_key("void kfr_amd::p2::k_pow2<float, 8, false, 3, 0>(kfr_amd::p2::args<float>)") == \
        "k_pow2<float, 8, false, 3, 0>"
    assert rs.kernel_key("void kfr_amd::(anonymous namespace)::k_fft<double, true, 2, false>(kfr_amd::lds_geom)") == \
        "k_fft<double, true, 2, false>"
    assert rs.kernel_key("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"


def test_traffic_table_is_read_for_the_named_kernel(tmp_path):
    f = tmp_path / "t.md"
    f.write_text("| dispatches | a | b | traffic bytes/dispatch | kernel |\n|---:|---:|---:|---:|---|\n"
                 "| 4 | 1.0 | 2.0 | 123456 | `k_pow2<float, 8, false, 3, 0>` |\n")
    assert rs.read_traffic(str(f), "k_pow2<float, 8, false, 3, 0>") == 123456.0
    assert rs.read_traffic(str(f), "k_pow2<float, 9, false, 3, 0>") is None
    assert rs.read_traffic(str(tmp_path / "missing.md"), "x") is None


def test_committed_traffic_table_names_a_kernel():
    table = ROOT / "profiles" / bench.TRAFFIC_TABLE
    rows = [l for l in table.read_text().splitlines() if l.startswith("| ") and "`k_" in l]
    assert rows, "the traffic summary lists no kernels"
    key = rows[0].split("`")[1]
    assert rs.read_traffic(str(table), key) > 0
