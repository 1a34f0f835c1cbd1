/* Test infrastructure: CPU restatement of the reference DFT path, in long double. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the checker; the product never
 * links it. Complex values are interleaved (re, im) double pairs. */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* reference_dft (include/kfr/dft/reference_dft.hpp:43-136): radix-2 recursion for powers of two,
 * direct O(n^2) sums otherwise; angles reduced exactly in integers. Unscaled in both directions. */
void oracle_dft(size_t n, const double* in, double* out, int inverse);

/* dft_plan_real forward (include/kfr/dft/fft.hpp:459-566, docs/docs/dft_format.md): spectrum of n
 * real values in CCs (fmt 1: n/2+1 values) or Perm (fmt 0: n/2 values, Nyquist real part in X[0].im). */
void oracle_rdft(size_t n, const double* in, double* out, int fmt);
/* dft_plan_real inverse: the unscaled inverse DFT of the Hermitian spectrum the packed values
 * describe, n real outputs. */
void oracle_irdft(size_t n, const double* in, double* out, int fmt);

/* dct_plan (include/kfr/dft/fft.hpp:927-970): DCT-II y_k = sum_n x_n cos(pi k (2n+1) / 2n_) and its
 * "inverse" DCT-III y_n = x_0/2 + sum_{k>0} x_k cos(pi k (2n+1) / 2n_), both unscaled. */
void oracle_dct(size_t n, const double* in, double* out, int inverse);

/* convolve / correlate (src/dft/convolution-impl.cpp:106-128) as direct sums: na+nb-1 outputs.
 * correlate uses b reversed (no conjugation). cplx: interleaved complex inputs; the reference keeps
 * only the real part of a complex result, which keep_real_only reproduces. */
void oracle_convolve(const double* a, size_t na, const double* b, size_t nb, double* out, int correlate,
                     int cplx, int keep_real_only);

/* Bluestein chirp of dft_arblen_stage_impl (src/dft/dft-impl.hpp:222-233): c[m] =
 * exp(-i pi k_m / n) with k_m accumulated as k += 2m-1 (mod 2n). */
void oracle_chirp(size_t n, double* out);

/* Permutations of the reference reorder stages, out[i] = in[perm[i]]:
 * fft_reorder (src/dft/bitrev.hpp:350-409): bit reversal of log2(n) bits;
 * dft_reorder_stage_impl (src/dft/dft-impl.hpp:428-494): digit reversal for the ascending radix list. */
void oracle_bitrev(size_t n, int64_t* perm);
void oracle_digitrev(const int* radices, size_t count, int64_t* perm);

#ifdef __cplusplus
}
#endif
