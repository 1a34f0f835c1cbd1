// Multidimensional DFTs (dft_plan_md / dft_plan_md_real, include/kfr/dft/fft.hpp:569-924) of
// row-major arrays on device data.
#pragma once

#include "fft_engine.hpp"
#include "fft_pow2.hpp"

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace kfr_amd
{

// Complex DFT over every axis of a row-major array of the given shape (two or more axes).
template <typename T>
class gpu_dft_md
{
public:
    explicit gpu_dft_md(const std::vector<size_t>& shape);
    ~gpu_dft_md();
    size_t size() const; // number of complex values
    const std::vector<size_t>& shape() const { return sh; }
    // one array; out == in allowed; no scaling
    void execute(void* out, const void* in, bool inverse, hipStream_t s);
    std::string schedule() const;

private:
    std::vector<size_t> sh;
    size_t total = 0;
    std::map<size_t, std::unique_ptr<gpu_dft<T>>> plans; // one per distinct axis length
    // power-of-two shapes: every axis but the last transformed in place along its stride
    bool strided = false;
    std::map<size_t, std::unique_ptr<pow2_fft<T>>> axis;
    dev_mem scratch;
};

// Real DFT over every axis: real array of `shape` <-> complex array of the same shape with the last
// axis n/2+1 (the CCs format of the last axis, which must be even).
template <typename T>
class gpu_dft_md_real
{
public:
    explicit gpu_dft_md_real(const std::vector<size_t>& shape);
    ~gpu_dft_md_real();
    size_t size() const;         // real values
    size_t complex_size() const; // complex values
    const std::vector<size_t>& shape() const { return sh; }
    void forward(void* cout, const void* rin, hipStream_t s);
    void inverse(void* rout, const void* cin, hipStream_t s);
    std::string schedule() const;

private:
    std::vector<size_t> sh;
    size_t rtotal = 0, ctotal = 0;
    std::map<size_t, std::unique_ptr<gpu_dft<T>>> plans; // complex axes
    std::unique_ptr<gpu_dft_real<T>> last;               // real last axis
    dev_mem a, b;
};

// out[c*rows + r] = in[r*cols + c] (complex values), LDS-tiled
template <typename T>
void transpose(void* out, const void* in, size_t rows, size_t cols, hipStream_t s);

extern template class gpu_dft_md<float>;
extern template class gpu_dft_md<double>;
extern template class gpu_dft_md_real<float>;
extern template class gpu_dft_md_real<double>;

} // namespace kfr_amd
