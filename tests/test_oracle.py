"""Pins the oracle (CPU): the long double restatement against the reference's known-answer vectors,
the golden outputs of the reference engine, the live reference build when present, and closed forms
that a plausible mistake (a wrong sign, index or conjugation) would fail."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from oracle import restated

GOLD = Path(__file__).resolve().parent / "golden"
KAT = json.loads((GOLD / "kat.json").read_text())
SP = np.load(GOLD / "spectra.npz")


def test_dct_known_answers():
    x = np.arange(16.0)
    assert np.sqrt(np.mean((restated.dct(x) - KAT["dct_fwd"]) ** 2)) < 1e-5
    assert np.sqrt(np.mean((restated.dct(x, True) - KAT["dct_inv"]) ** 2)) < 1e-5


def test_convolution_known_answers():
    a = np.array(KAT["conv_a"], dtype=np.float64)
    b = np.array(KAT["conv_b"], dtype=np.float64)
    assert np.allclose(restated.convolve(a, b), KAT["convolve"], atol=1e-12)
    assert np.allclose(restated.convolve(a, b, correlate=True), KAT["correlate"], atol=1e-12)
    c = restated.convolve(a.astype(np.complex128), b.astype(np.complex128))
    assert np.allclose(c, KAT["convolve"], atol=1e-12)


def test_capi_round_trip_known_answer():
    n = 32
    x = np.array([complex(i / 32, -i / 32) for i in range(n)])
    z = restated.dft(restated.dft(x), inverse=True)
    assert np.max(np.abs(z - np.array([complex(i, -i) for i in range(n)]))) < 1e-12


@pytest.mark.parametrize("tn,tol", [("f32", 4e-6), ("f64", 1e-13)])
def test_against_reference_golden_spectra(tn, tol):
    for n in [1, 2, 3, 5, 8, 16, 48, 60, 64, 97, 101, 127, 210, 1024, 4096]:
        x = SP[f"c_in_{tn}_{n}"]
        for inv, key in ((False, "c_fwd"), (True, "c_inv")):
            e = oracle.rel_rms(SP[f"{key}_{tn}_{n}"], restated.dft(x, inv))
            assert e < tol, (n, inv, e)
    for n in [2, 4, 30, 64, 2048]:
        x = SP[f"r_in_{tn}_{n}"]
        for fmt in (0, 1):
            X = restated.rdft(x, fmt)
            assert oracle.rel_rms(SP[f"r_fwd_{tn}_{n}_{fmt}"], X) < tol, (n, fmt)
            assert oracle.rel_rms(SP[f"r_inv_{tn}_{n}_{fmt}"], restated.irdft(SP[f"r_fwd_{tn}_{n}_{fmt}"], n, fmt)) < tol
    for n in [16, 17, 64]:
        x = SP[f"dct_in_{tn}_{n}"]
        assert oracle.rel_rms(SP[f"dct_fwd_{tn}_{n}"], restated.dct(x)) < tol
        assert oracle.rel_rms(SP[f"dct_inv_{tn}_{n}"], restated.dct(x, True)) < tol
    a, b = SP[f"conv_a_{tn}"], SP[f"conv_b_{tn}"]
    assert oracle.rel_rms(SP[f"conv_{tn}"], restated.convolve(a, b)) < tol
    assert oracle.rel_rms(SP[f"corr_{tn}"], restated.convolve(a, b, correlate=True)) < tol
    ac = restated.convolve(a, a, correlate=True)
    assert oracle.rel_rms(SP[f"acorr_{tn}"], ac[ac.size // 2:]) < tol
    taps, sig = SP[f"filt_taps_{tn}"], SP[f"filt_in_{tn}"]
    direct = restated.convolve(sig, taps)[: sig.size]
    assert oracle.rel_rms(SP[f"filt_out_{tn}"], direct) < tol * 10


def test_complex_convolve_drops_imaginary_part():
    """the reference returns real(...) for complex convolution (src/dft/convolution-impl.cpp:124-127)"""
    a, b = SP["cconv_a"], SP["cconv_b"]
    ours = restated.convolve(a, b, keep_real_only=True)
    assert np.allclose(ours.imag, 0)
    assert oracle.rel_rms(SP["cconv"], ours) < 1e-6
    assert oracle.rel_rms(SP["cconv"], restated.convolve(a, b, keep_real_only=False)) > 0.1


def test_against_live_reference():
    ref = oracle.reference()
    if ref is None:
        pytest.skip("reference build not present")
    for n in [2, 7, 16, 100, 128, 243, 1000, 2048, 3150, 4096, 9990, 211]:
        for cdt in (np.complex64, np.complex128):
            x = oracle.complex_gaussian(n, n, cdt)
            for inv in (False, True):
                tol = 4e-6 if cdt == np.complex64 else 1e-13
                assert oracle.rel_rms(ref.dft(x, inv), restated.dft(x, inv)) < tol, (n, cdt, inv)


# closed forms and invariants, each chosen so that a sign, index or conjugation slip fails


def test_closed_forms():
    n = 64
    imp = np.zeros(n, complex)
    imp[0] = 1
    assert np.allclose(restated.dft(imp), 1)                              # impulse -> ones
    imp[3] = 0
    imp[0] = 0
    imp[1] = 1
    k = np.arange(n)
    assert np.allclose(restated.dft(imp), np.exp(-2j * np.pi * k / n))    # shifted impulse -> phase ramp (sign)
    assert np.allclose(restated.dft(imp, True), np.exp(2j * np.pi * k / n))
    x = oracle.complex_gaussian(1, n, np.complex128)
    y = oracle.complex_gaussian(2, n, np.complex128)
    X = restated.dft(x)
    assert np.isclose(np.sum(np.abs(X) ** 2), n * np.sum(np.abs(x) ** 2))  # Parseval
    assert np.allclose(restated.dft(2 * x + 3j * y), 2 * X + 3j * restated.dft(y))  # linearity
    assert np.allclose(restated.dft(X, True) / n, x)                          # round trip
    # DFT matrix by brute force at n = 6
    m = 6
    W = np.exp(-2j * np.pi * np.outer(np.arange(m), np.arange(m)) / m)
    v = oracle.complex_gaussian(3, m, np.complex128)
    assert np.allclose(restated.dft(v), W @ v)
    # real transform: Hermitian output of real input
    r = oracle.gaussian(4, 30, np.float64)
    full = restated.dft(r.astype(complex))
    assert np.allclose(restated.rdft(r, 1), full[:16])
    assert np.allclose(restated.irdft(restated.rdft(r, 0), 30, 0) / 30, r)
    # DCT against the cosine matrix
    d = oracle.gaussian(5, 9, np.float64)
    C = np.cos(np.pi * np.outer(np.arange(9), 2 * np.arange(9) + 1) / 18)
    assert np.allclose(restated.dct(d), C @ d)
    assert np.allclose(restated.dct(restated.dct(d), True), 9 / 2 * d)      # DCT-III(DCT-II x) = N/2 x
    # convolution against numpy
    a, b = oracle.gaussian(6, 17, np.float64), oracle.gaussian(7, 5, np.float64)
    assert np.allclose(restated.convolve(a, b), np.convolve(a, b))
    assert np.allclose(restated.convolve(a, b, correlate=True), np.convolve(a, b[::-1]))
    # Bluestein chirp: the recurrence of the reference equals the closed form m^2 mod 2n
    for nn in [101, 127, 3163]:
        mm = np.arange(nn)
        assert np.allclose(restated.chirp(nn), np.exp(-1j * np.pi * ((mm * mm) % (2 * nn)) / nn))


def test_planted_mistake_is_caught():
    """A conjugated spectrum (the classic sign slip) must fail the same bound the GPU tests use."""
    x = SP["c_in_f32_210"]
    good = restated.dft(x)
    bad = np.conj(restated.dft(np.conj(x), True))  # equals good: conj(IDFT(conj x)) = DFT(x)
    assert oracle.rel_rms(bad, good) < 1e-12
    wrong = restated.dft(x, True)                  # the slip: inverse sign used for the forward transform
    assert oracle.rel_rms(SP["c_fwd_f32_210"], wrong) > 0.5


def test_generator_is_deterministic():
    a = oracle.uniform(3, 1000)
    assert np.array_equal(a, oracle.uniform(3, 1000))
    assert a.min() >= -1 and a.max() < 1
    g = oracle.gaussian(4, 100000, np.float64)
    assert abs(g.mean()) < 0.02 and abs(g.std() - 1) < 0.02
