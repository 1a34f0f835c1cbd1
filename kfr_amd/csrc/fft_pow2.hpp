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
    // true for N = 2^L that one kernel (L <= 14 for f32, 13 for f64), the two-kernel four-step split
    // (up to twice that) or the multi-pass Stockham over HBM (the longest lengths) covers
    static bool supports(size_t n);
    // prec: storage of the data in HBM, 0 = complex T, 1 = bfloat16 pairs, 2 = binary16 pairs (T = float,
    // radix-16 passes on MFMA)
    pow2_fft(size_t n, hipStream_t s, int prec = 0);
    ~pow2_fft();
    // batch transforms: in and out hold batch*n complex values each, item b at offset b*n; in == out
    // is allowed. No scaling (the inverse is unnormalised).
    void run(bool inverse, const void* in, void* out, long long batch, hipStream_t s);
    std::string describe() const;

private:
    struct impl;
    std::unique_ptr<impl> p;
};

extern template class pow2_fft<float>;
extern template class pow2_fft<double>;

} // namespace kfr_amd
