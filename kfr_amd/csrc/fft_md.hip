// Multidimensional DFTs on device data. Like the reference (dft_plan_md::execute_dim,
// include/kfr/dft/fft.hpp:641-690), every axis is transformed as contiguous rows followed by a
// transpose that rotates the next axis to the end, so each axis reuses the batched 1-D engine (and
// its power-of-two fast path). Here each transform reads the previous step's array and writes a
// scratch array that the LDS-tiled transpose turns back into the destination, so every step is one
// HBM round trip for the transform and one for the transpose.
#include "fft_md.hpp"

#include "fft_core.hpp"

#include <stdexcept>

namespace kfr_amd
{

// 32x32 tiles through LDS (padded row), 256 threads, each moving 4 values in and 4 out
template <typename T>
__global__ void __launch_bounds__(256) k_transpose(const cx<T>* __restrict__ in, cx<T>* __restrict__ out,
                                                   long long rows, long long cols, long long tiles_x)
{
    __shared__ cx<T> tile[32][33];
    const long long t  = blockIdx.x;
    const long long ty = t / tiles_x;
    const long long tx = t - ty * tiles_x;
    const int lx       = threadIdx.x & 31;
    const int ly       = threadIdx.x >> 5;
    const long long c  = tx * 32 + lx;
#pragma unroll
    for (int k = 0; k < 32; k += 8)
    {
        const long long r = ty * 32 + ly + k;
        if (r < rows && c < cols)
            tile[ly + k][lx] = in[r * cols + c];
    }
    __syncthreads();
    const long long r2 = ty * 32 + lx;
#pragma unroll
    for (int k = 0; k < 32; k += 8)
    {
        const long long c2 = tx * 32 + ly + k;
        if (c2 < cols && r2 < rows)
            out[c2 * rows + r2] = tile[lx][ly + k];
    }
}

template <typename T>
void transpose(void* out, const void* in, size_t rows, size_t cols, hipStream_t s)
{
    if (rows == 0 || cols == 0)
        return;
    const long long tx = (long long)((cols + 31) / 32), ty = (long long)((rows + 31) / 32);
    if (tx * ty > 0xffffffffll)
        throw std::runtime_error("array too large to transpose in one launch");
    hipLaunchKernelGGL(k_transpose<T>, dim3((unsigned)(tx * ty)), dim3(256), 0, s, static_cast<const cx<T>*>(in),
                       static_cast<cx<T>*>(out), (long long)rows, (long long)cols, tx);
    hip_check(hipGetLastError(), "transpose launch");
}
template void transpose<float>(void*, const void*, size_t, size_t, hipStream_t);
template void transpose<double>(void*, const void*, size_t, size_t, hipStream_t);

namespace
{
size_t product(const std::vector<size_t>& v)
{
    size_t p = 1;
    for (size_t x : v)
        p *= x;
    return p;
}
void check_shape(const std::vector<size_t>& sh)
{
    if (sh.size() < 2)
        throw std::invalid_argument("a multidimensional plan needs at least two axes");
    for (size_t x : sh)
        if (x == 0)
            throw std::invalid_argument("every axis of a multidimensional plan must be positive");
    size_t p = 1;
    for (size_t x : sh)
    {
        if (p > ((size_t)1 << 40) / x)
            throw std::invalid_argument("multidimensional plan is too large");
        p *= x;
    }
}
template <typename T>
std::string axis_text(const std::map<size_t, std::unique_ptr<gpu_dft<T>>>& plans, const std::vector<size_t>& sh,
                      size_t axes)
{
    std::string s;
    for (size_t i = 0; i < axes; ++i)
    {
        const size_t n = sh[i];
        s += "axis " + std::to_string(i) + " (" + std::to_string(n) + "): ";
        auto it = plans.find(n);
        s += (n > 1 && it != plans.end()) ? it->second->schedule().text() : std::string("identity");
        s += "\n";
    }
    return s;
}
} // namespace

template <typename T>
gpu_dft_md<T>::gpu_dft_md(const std::vector<size_t>& shape) : sh(shape)
{
    check_shape(sh);
    total = product(sh);
    // power-of-two axes of one-kernel length: the axes before the last run in place along their
    // stride (one HBM round trip per axis, no transposes)
    strided = (sh.back() & (sh.back() - 1)) == 0;
    for (size_t i = 0; i + 1 < sh.size(); ++i)
        strided = strided && (sh[i] == 1 || pow2_fft<T>::supports_axis(sh[i]));
    if (strided)
    {
        if (sh.back() > 1)
            plans[sh.back()].reset(new gpu_dft<T>(sh.back()));
        hipStream_t s = library_stream();
        for (size_t i = 0; i + 1 < sh.size(); ++i)
            if (sh[i] > 1 && !axis.count(sh[i]))
                axis[sh[i]].reset(new pow2_fft<T>(sh[i], s));
        hip_check(hipStreamSynchronize(s), "multidimensional plan initialisation");
        return;
    }
    for (size_t n : sh)
        if (n > 1 && !plans.count(n))
            plans[n].reset(new gpu_dft<T>(n));
}
template <typename T>
gpu_dft_md<T>::~gpu_dft_md() = default;

template <typename T>
size_t gpu_dft_md<T>::size() const
{
    return total;
}

template <typename T>
std::string gpu_dft_md<T>::schedule() const
{
    if (strided)
    {
        std::string t;
        for (size_t i = 0; i + 1 < sh.size(); ++i)
            t += "axis " + std::to_string(i) + " (" + std::to_string(sh[i]) + "): " +
                 (sh[i] > 1 ? "in place along its stride, " + axis.at(sh[i])->describe() : std::string("identity")) +
                 "\n";
        const size_t n = sh.size() - 1;
        return t + "axis " + std::to_string(n) + " (" + std::to_string(sh.back()) + "): " +
               (sh.back() > 1 ? plans.at(sh.back())->schedule().text() : std::string("identity"));
    }
    return axis_text(plans, sh, sh.size()) + "transpose after every axis";
}

template <typename T>
void gpu_dft_md<T>::execute(void* out, const void* in, bool inverse, hipStream_t s)
{
    if (strided)
    {
        const size_t last = sh.back();
        if (last > 1)
            plans.at(last)->execute(out, in, (long long)(total / last), inverse, s);
        else if (out != in)
            hip_check(hipMemcpyAsync(out, in, total * sizeof(cx<T>), hipMemcpyDeviceToDevice, s), "copy");
        int inner_log = 0;
        for (size_t j = sh.size() - 1; j-- > 0;)
        {
            size_t t = sh[j + 1];
            while (t > 1)
            {
                ++inner_log;
                t >>= 1;
            }
            if (sh[j] > 1)
            {
                long long outer = 1;
                for (size_t q = 0; q < j; ++q)
                    outer *= (long long)sh[q];
                axis.at(sh[j])->run_axis(inverse, out, out, inner_log, outer, s);
            }
        }
        return;
    }
    scratch.ensure(total * sizeof(cx<T>));
    std::vector<size_t> cur = sh; // current layout, last axis contiguous
    const void* src         = in;
    for (size_t k = 0; k < sh.size(); ++k)
    {
        const size_t len  = cur.back();
        const size_t rows = total / len;
        if (len > 1)
            plans.at(len)->execute(scratch.ptr, src, (long long)rows, inverse, s);
        else
            hip_check(hipMemcpyAsync(scratch.ptr, src, total * sizeof(cx<T>), hipMemcpyDeviceToDevice, s), "copy");
        transpose<T>(out, scratch.ptr, rows, len, s);
        // the transposed array has the transformed axis first
        cur.insert(cur.begin(), cur.back());
        cur.pop_back();
        src = out;
    }
}

template <typename T>
gpu_dft_md_real<T>::gpu_dft_md_real(const std::vector<size_t>& shape) : sh(shape)
{
    check_shape(sh);
    rtotal = product(sh);
    last.reset(new gpu_dft_real<T>(sh.back(), 1 /* CCs */));
    ctotal = rtotal / sh.back() * last->complex_size();
    for (size_t i = 0; i + 1 < sh.size(); ++i)
        if (sh[i] > 1 && !plans.count(sh[i]))
            plans[sh[i]].reset(new gpu_dft<T>(sh[i]));
}
template <typename T>
gpu_dft_md_real<T>::~gpu_dft_md_real() = default;

template <typename T>
size_t gpu_dft_md_real<T>::size() const
{
    return rtotal;
}
template <typename T>
size_t gpu_dft_md_real<T>::complex_size() const
{
    return ctotal;
}

template <typename T>
std::string gpu_dft_md_real<T>::schedule() const
{
    return axis_text(plans, sh, sh.size() - 1) + "axis " + std::to_string(sh.size() - 1) + " (" +
           std::to_string(sh.back()) + ", real): " + last->schedule().text() + "\ntranspose after every axis";
}

template <typename T>
void gpu_dft_md_real<T>::forward(void* cout, const void* rin, hipStream_t s)
{
    a.ensure(ctotal * sizeof(cx<T>));
    b.ensure(ctotal * sizeof(cx<T>));
    const size_t ck    = last->complex_size();
    const size_t rows0 = rtotal / sh.back();
    last->forward(a.ptr, rin, (long long)rows0, s);
    std::vector<size_t> cur(sh.begin(), sh.end() - 1);
    cur.insert(cur.begin(), ck); // after the transpose: (ck, d0, ..., d_{k-1})
    const size_t naxes = sh.size() - 1;
    transpose<T>(naxes == 0 ? cout : b.ptr, a.ptr, rows0, ck, s);
    for (size_t k = 0; k < naxes; ++k)
    {
        const size_t len  = cur.back();
        const size_t rows = ctotal / len;
        if (len > 1)
            plans.at(len)->execute(a.ptr, b.ptr, (long long)rows, false, s);
        else
            hip_check(hipMemcpyAsync(a.ptr, b.ptr, ctotal * sizeof(cx<T>), hipMemcpyDeviceToDevice, s), "copy");
        transpose<T>(k + 1 == naxes ? cout : b.ptr, a.ptr, rows, len, s);
        cur.insert(cur.begin(), cur.back());
        cur.pop_back();
    }
}

template <typename T>
void gpu_dft_md_real<T>::inverse(void* rout, const void* cin, hipStream_t s)
{
    a.ensure(ctotal * sizeof(cx<T>));
    b.ensure(ctotal * sizeof(cx<T>));
    const size_t ck    = last->complex_size();
    const size_t naxes = sh.size() - 1;
    // layout (d0, ..., d_{k-1}, ck); bring each complex axis to the end, transform it in place
    std::vector<size_t> cur(sh.begin(), sh.end() - 1);
    cur.push_back(ck);
    const void* src = cin;
    void* bufs[2]   = { a.ptr, b.ptr };
    int which       = 0;
    for (size_t k = 0; k < naxes; ++k)
    {
        const size_t front = cur.front();
        const size_t rest  = ctotal / front;
        void* dst          = bufs[which];
        transpose<T>(dst, src, front, rest, s);
        if (front > 1)
            plans.at(front)->execute(dst, dst, (long long)rest, true, s);
        cur.push_back(cur.front());
        cur.erase(cur.begin());
        src = dst;
        which ^= 1;
    }
    // layout (ck, d0, ..., d_{k-1}): back to (d0, ..., ck), then the real inverse of every row
    void* dst = bufs[which];
    transpose<T>(dst, src, ck, ctotal / ck, s);
    last->inverse(rout, dst, (long long)(rtotal / sh.back()), s);
}

template class gpu_dft_md<float>;
template class gpu_dft_md<double>;
template class gpu_dft_md_real<float>;
template class gpu_dft_md_real<double>;

} // namespace kfr_amd
