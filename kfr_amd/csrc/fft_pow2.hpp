// Power-of-two complex transforms over contiguous batches: the plain C2C execute path of
// dft_plan<T> (include/kfr/dft/fft.hpp:183-456) for N = 2^L, built from kernels specialised at
// compile time for each L.
#pragma once

#include "fft_engine.hpp"

#include <memory>
#include <string>

namespace kfr_amd
{

template <typename T>
class pow2_fft
{
public:
    // true for N = 2^L that one kernel (L <= 15 for f32, 14 for f64), the two-kernel four-step split
    // (up to twice that) or the multi-pass Stockham over HBM (the longest lengths) covers
    static bool supports(size_t n);
    // prec: storage of the data in HBM, 0 = complex T, 1 = bfloat16 pairs, 2 = binary16 pairs (T = float,
    // radix-16 passes on MFMA)
    pow2_fft(size_t n, hipStream_t s, int prec = 0);
    ~pow2_fft();
    // batch transforms: in and out hold batch*n complex values each, item b at offset b*n; in == out
    // is allowed. No scaling (the inverse is unnormalised).
    void run(bool inverse, const void* in, void* out, long long batch, hipStream_t s, workspace w = {});
    // lengths whose transform along a strided axis runs in one kernel (2^7 .. the one-kernel maximum)
    static bool supports_axis(size_t n);
    // transforms along a strided axis, in == out allowed: element (o, i, c) at ((o*n + i) << inner_log) + c
    // for o < outer, i < n, c < 2^inner_log (full precision, one-kernel lengths only)
    void run_axis(bool inverse, const void* in, void* out, int inner_log, long long outer, hipStream_t s);
    // DCT-II (dct_plan, unscaled) of n = 2N reals for this plan's length N in one kernel: the mirrored
    // input packed as N complex values on load, the N-point transform, and the real spectrum formed
    // on store; inverse: DCT-III (dct_plan's inverse, unscaled), the same steps backwards. wn holds
    // W_n^k (k <= N) and w4n W_{4n}^k (k < n). False when this length has no fused kernel (not one
    // workgroup, under 16 threads per transform, or half-precision storage).
    bool dct(bool inverse, const void* x, void* y, const void* wn, const void* w4n, long long batch, hipStream_t s);
    // real transform (dft_plan_real) of n = 2N reals for this plan's length N in one kernel: forward
    // reads the reals as N complex values and writes the CCs (fmt 1, N + 1 values) or Perm (fmt 0)
    // spectrum at distance odist; inverse reads spectra at distance idist and writes the reals. rtw
    // holds exp(-i(2 pi k/n + pi/2)) for k <= N/2. False when this length has no fused kernel.
    bool real(bool inverse, const void* in, long long idist, void* out, long long odist, const void* rtw, int fmt,
              long long batch, hipStream_t s);
    std::string describe() const;

private:
    struct impl;
    std::unique_ptr<impl> p;
};

extern template class pow2_fft<float>;
extern template class pow2_fft<double>;

// Batched linear convolution of real signals with a short filter by overlap-save: one kernel per call
// transforms blocks of 2^L points (two segments per complex transform), multiplies by the filter
// spectrum and transforms back in LDS. Same contract as gpu_conv (convolve / correlate and the
// adjoint); available when the filter fits an LDS block with room for the valid outputs.
template <typename T>
class os_conv
{
public:
    // log2 of the block length for these lengths, or -1 when the filter is too long
    static int block_log(size_t nx, size_t nh);
    os_conv(size_t nx, size_t nh, hipStream_t s);
    ~os_conv();
    // y[b] = x[b] * h[b] (h_dist = 0: one shared filter); correlate: h reversed. |y| = nx + nh - 1
    void forward(T* y, const T* x, const T* h, long long h_dist, long long batch, bool correlate, hipStream_t s);
    // adjoint of forward: dx = g correlated with the filter (h, or h reversed when correlate), and the
    // filter gradient g correlated with x (first nh lags; summed over the batch when h_dist = 0) into dh,
    // reversed when correlate is set by the caller (gpu_conv); either output may be null
    void backward(T* dx, T* dh, const T* g, const T* x, const T* h, long long h_dist, long long batch,
                  bool correlate, hipStream_t s);
    std::string describe() const;

    struct impl;

private:
    std::unique_ptr<impl> p;
};

extern template class os_conv<float>;
extern template class os_conv<double>;

} // namespace kfr_amd
