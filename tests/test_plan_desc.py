"""Plan-structure parity (CPU): the library's reference plan descriptor and reorder permutations
against the reference's own plans (golden fixtures, and the live reference build when present)."""
from __future__ import annotations

import json
import re
from pathlib import Path

import numpy as np
import pytest

import kfr_amd
import oracle

GOLD = Path(__file__).resolve().parent / "golden"
PLANS = json.loads((GOLD / "plans.json").read_text())


def _strip(s: str) -> str:
    return re.sub(r"\((avx2|avx512|sse2|sse41|avx|gfx950)\)", "", s)


def _ours(key: str) -> str:
    parts = key.split(":")
    kind, t, n = parts[0], parts[1], int(parts[2])
    if kind == "dft":
        return _strip(kfr_amd.describe(0, t, n))
    if kind == "dft_progressive":
        return _strip(kfr_amd.describe(3, t, n))
    if kind == "real":
        return _strip(kfr_amd.describe(1, t, n, int(parts[3])))
    return _strip(kfr_amd.describe(2, t, n))


@pytest.mark.parametrize("key", sorted(k for k in PLANS if k != "reference"))
def test_descriptor_matches_reference_fixture(key):
    assert _ours(key) == PLANS[key], key


def test_descriptor_against_live_reference():
    ref = oracle.reference()
    if ref is None:
        pytest.skip("reference build not present")
    rng = np.random.default_rng(5)
    sizes = sorted(set(list(range(1, 600)) + [int(v) for v in rng.integers(600, 200000, 200)] +
                   [int(v) for v in rng.integers(1, 40, 50) * 1024]))
    bad = []
    for t, dt in (("f32", 0), ("f64", 1)):
        for n in sizes:
            if _strip(kfr_amd.describe(0, t, n)) != _strip(ref.describe(0, dt, n)):
                bad.append((t, n))
            if n % 2 == 0 and n < 100000:
                for fmt in (0, 1):
                    if _strip(kfr_amd.describe(1, t, n, fmt)) != _strip(ref.describe(1, dt, n, fmt)):
                        bad.append((t, n, "real", fmt))
    assert not bad, bad[:20]


def test_dump_layout():
    text = kfr_amd.describe(0, "f32", 4096)
    lines = text.splitlines()
    assert lines[0] == "size=4096 temp_size=32832 data_size=32768 stages=6,6 arblen=0"
    assert lines[1] == "fft_autosort_stage_impl<float, true, false, false>(gfx950): 4, 16384, 24576, 0, 1, 0, 4096, 0, 0"


PERMS = np.load(GOLD / "perms.npz")


@pytest.mark.parametrize("key", sorted(PERMS.files))
def test_permutation_matches_reference_fixture(key):
    t, n = key.split("_")
    p = kfr_amd.permutation(t, int(n))
    assert p is not None
    assert np.array_equal(p, PERMS[key].astype(np.int64))


def _radices(t: str, n: int):
    """init_dft's ascending radix list, read from the descriptor's fixed/generic stages"""
    radices = []
    for line in kfr_amd.describe(0, t, n).splitlines()[1:]:
        m = re.match(r"dft_stage_(fixed|fixed_final|generic)_impl<[^>]*>\([^)]*\): (\d+),", line)
        if m:
            radices.append(int(m.group(2)))
    return radices


@pytest.mark.parametrize("n", [32768, 65536, 1 << 18])
def test_bitrev_closed_form(n):
    p = kfr_amd.permutation("f32", n)
    assert np.array_equal(p, oracle.restated.bitrev(n))


@pytest.mark.parametrize("n", [210, 1029, 3150, 8000, 9990, 360, 120, 30720000 // 2 // 100, 7 * 7 * 7 * 9 * 10])
def test_digitrev_closed_form(n):
    p = kfr_amd.permutation("f32", n)
    if p is None:
        pytest.skip("plan has two or fewer stages")
    assert np.array_equal(p, oracle.restated.digitrev(_radices("f32", n)))


def test_permutations_are_bijections():
    for n in [4096, 1 << 15, 210, 9990, 1 << 20]:
        for t in ("f32", "f64"):
            p = kfr_amd.permutation(t, n)
            if p is not None:
                assert np.array_equal(np.sort(p), np.arange(n))


def test_progressive_total_steps():
    assert kfr_amd.lib().kfr_amd_dft_progressive_total_steps(0, 4096) == 6
    assert kfr_amd.lib().kfr_amd_dft_progressive_total_steps(0, 1 << 24) == 9
    assert kfr_amd.lib().kfr_amd_dft_progressive_total_steps(1, 4096) == 3


def test_real_plan_requires_even_size():
    L = kfr_amd.lib()
    h = L.kfr_dft_real_create_plan_f32(7, 1)
    assert not h
    assert "even" in kfr_amd.last_error()
    # a successful call clears the error
    assert kfr_amd.describe(0, "f32", 16)
    assert kfr_amd.last_error() == ""
