// Mixed-radix complex transforms over contiguous batches with kernels specialised at compile time for
// a menu of lengths (2^a 3^b 5^c 7^d between 1000 and 4200), used directly or as both factors of the
// four-step split of longer lengths. The plain C2C execute path of dft_plan<T>
// (include/kfr/dft/fft.hpp:183-456) for those lengths.
#pragma once

#include "fft_engine.hpp"

#include <memory>
#include <string>

namespace kfr_amd
{

template <typename T>
class mixed_fft
{
public:
    // true when n is a menu length, or n = n1*n2 with both factors menu lengths (T = float)
    static bool supports(size_t n);
    mixed_fft(size_t n, hipStream_t s);
    ~mixed_fft();
    // batch transforms of contiguous items (distance n), in == out allowed, unscaled
    void run(bool inverse, const void* in, void* out, long long batch, hipStream_t s, workspace w = {});
    std::string describe() const;

private:
    struct impl;
    std::unique_ptr<impl> p;
};

extern template class mixed_fft<float>;
extern template class mixed_fft<double>;

} // namespace kfr_amd
