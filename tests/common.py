"""Helpers shared by the parity tests: FP64 references and tolerances."""
from __future__ import annotations

import math

import numpy as np

import oracle

EPS = {np.dtype(np.complex64): float(np.finfo(np.float32).eps), np.dtype(np.complex128): float(np.finfo(np.float64).eps),
       np.dtype(np.float32): float(np.finfo(np.float32).eps), np.dtype(np.float64): float(np.finfo(np.float64).eps)}


def tol_kfr(n: int, dtype) -> float:
    """Spectra parity bound against the reference engine's own output: relative RMS error
    4*log2(N)*eps of the stated dtype."""
    return 4.0 * max(1.0, math.log2(max(n, 2))) * EPS[np.dtype(dtype)]


def bound_fp64(n: int, dtype) -> float:
    """Bound on the relative RMS error against an FP64 (or better) DFT: 1e-6 for single precision at
    every length (the accuracy target of the single-precision path), 4*log2(N)*eps but at least 1e-14
    for double precision."""
    if np.dtype(dtype) in (np.dtype(np.complex64), np.dtype(np.float32)):
        return 1e-6
    return max(1e-14, tol_kfr(n, dtype))


def fp64_dft(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """Unscaled DFT in FP64 or better: the long double oracle where it is fast, numpy's FP64 FFT beyond."""
    n = x.shape[-1]
    pow2 = n & (n - 1) == 0
    if (pow2 and n <= 1 << 16) or n <= 4096:
        x2 = np.ascontiguousarray(x, dtype=np.complex128).reshape(-1, n)
        return np.stack([oracle.restated.dft(r, inverse) for r in x2]).reshape(x.shape)
    xx = np.asarray(x, dtype=np.complex128)
    return np.fft.ifft(xx, axis=-1) * n if inverse else np.fft.fft(xx, axis=-1)


def fp64_rdft(x: np.ndarray, fmt: int = 1) -> np.ndarray:
    n = x.shape[-1]
    if n <= 4096 or (n & (n - 1) == 0 and n <= 1 << 16):
        x2 = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, n)
        return np.stack([oracle.restated.rdft(r, fmt) for r in x2]).reshape(x.shape[:-1] + (-1,))
    X = np.fft.rfft(np.asarray(x, dtype=np.float64), axis=-1)
    if fmt == 0:
        X = X.copy()
        X[..., 0] = X[..., 0].real + 1j * X[..., -1].real
        X = X[..., :-1]
    return X


def fp64_dct(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    n = x.shape[-1]
    if n <= 2048:
        x2 = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, n)
        return np.stack([oracle.restated.dct(r, inverse) for r in x2]).reshape(x.shape)
    import scipy.fft
    # scipy's unnormalised DCT-II is 2 sum x cos(...); DCT-III is x0 + 2 sum_{k>0} x_k cos(...)
    return scipy.fft.dct(np.asarray(x, dtype=np.float64), type=3 if inverse else 2, axis=-1) / 2.0


def uneven_lengths() -> np.ndarray:
    """Lengths of the mixed-length batch workload: every 7-smooth N in [18, 8192] (2^a 3^b 5^c 7^d)
    and every prime in [17, 127]."""
    smooth = set()
    for a in range(14):
        for b in range(9):
            for c in range(6):
                for d in range(5):
                    v = 2 ** a * 3 ** b * 5 ** c * 7 ** d
                    if 18 <= v <= 8192:
                        smooth.add(v)
    primes = [p for p in range(17, 128) if all(p % q for q in range(2, int(p ** 0.5) + 1))]
    return np.array(sorted(smooth | set(primes)), dtype=np.int64)
