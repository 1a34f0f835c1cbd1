// Host interface of the gfx950 DFT engine: GPU plans for complex, real and DCT transforms and the
// FFT convolution, all operating on device pointers on a HIP stream. The C ABI (capi.cpp) wraps
// these with the reference's host-pointer semantics.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <algorithm>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace kfr_amd
{

struct hip_error : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what);

// Test hook: when > 0, the n-th following device allocation made by the library fails (counted down
// on every allocation), so the error path of every entry point can be exercised from the host.
void set_alloc_failure_countdown(int n);

// Caller-provided device memory for a plan's intermediates (four-step and multi-pass rounds): when p
// is set, the plan uses it instead of allocating its own scratch and runs in rounds that fit `bytes`
// (rounds_in); a plan that needs intermediates beyond one item's worth there throws.
struct workspace
{
    void* p      = nullptr;
    size_t bytes = 0;
};

// batch items per round: `chunk`, or what `buffers` equal 256-byte aligned parts of the workspace hold
// at `per_item` bytes each
inline long long rounds_in(const workspace& w, long long chunk, size_t per_item, int buffers)
{
    if (!w.p)
        return chunk;
    const size_t each = (w.bytes / (size_t)buffers) & ~(size_t)255;
    const long long c = (long long)(each / per_item);
    if (c < 1)
        throw std::length_error("the workspace of " + std::to_string(w.bytes) + " bytes cannot hold the " +
                                std::to_string((size_t)buffers * per_item) + " bytes of one item's intermediate");
    return std::min(chunk, c);
}
// the b-th of `buffers` equal aligned parts of the workspace
inline void* workspace_part(const workspace& w, int b, int buffers)
{
    const size_t each = (w.bytes / (size_t)buffers) & ~(size_t)255;
    return static_cast<char*>(w.p) + (size_t)b * each;
}

// The workspace of the allocation-free entry point: a static device array of noalloc_scratch_bytes
// (fft_noalloc.hip), its address on the current device looked up on first use there. Calls that use
// it and overlap on one GPU race.
constexpr size_t noalloc_scratch_bytes = (size_t)32 << 20;
workspace noalloc_workspace();

// Owned device allocation
struct dev_mem
{
    void* ptr    = nullptr;
    size_t bytes = 0;
    dev_mem()    = default;
    explicit dev_mem(size_t b) { reset(b); }
    dev_mem(const dev_mem&)            = delete;
    dev_mem& operator=(const dev_mem&) = delete;
    ~dev_mem();
    void reset(size_t b);
    void ensure(size_t b)
    {
        if (b > bytes)
            reset(b);
    }
    template <typename U>
    U* as() const
    {
        return static_cast<U*>(ptr);
    }
};

// The GPU schedule chosen for a complex transform length, for dump() and tests.
struct schedule_info
{
    std::string algorithm;         // "lds", "four-step", "bluestein-fused", "bluestein"
    std::vector<int> radices;      // LDS passes (lds) or of each sub-transform
    size_t n1 = 0, n2 = 0;         // four-step split (rows x columns)
    size_t bluestein_m = 0;        // Bluestein FFT length
    int threads_per_transform = 0; // lds
    int transforms_per_block  = 0; // lds
    std::string contiguous;        // kernels that contiguous batches use instead, when they differ
    std::string text() const;
};

// Complex DFT of length n on device data (contiguous batches with distance n, in-place allowed).
template <typename T>
class gpu_dft
{
public:
    explicit gpu_dft(size_t n);
    ~gpu_dft();
    size_t size() const;
    const schedule_info& schedule() const;
    void execute(void* out, const void* in, long long batch, bool inverse, hipStream_t s);
    // general strided batches: element i of transform b at base + b*dist + i (complex units)
    void execute_dist(void* out, long long odist, const void* in, long long idist, long long batch,
                      bool inverse, hipStream_t s);
    // execute() with the intermediates in `w` (no allocation by the plan); Bluestein lengths whose
    // transform does not fit one workgroup are refused
    void execute_workspace(void* out, const void* in, long long batch, bool inverse, hipStream_t s, workspace w);
    // Progressive execution of one transform in place on device data (n values): progressive_prepare
    // lowers the transform to about `want` units (one launch each; Stockham passes of a finer schedule
    // for lengths in one workgroup, the column and the row kernel of a four-step, the whole transform
    // for Bluestein) and returns their count; progressive_unit(0 .. count-1) then run in order.
    // Four-step lengths keep their intermediate in `mid` (progressive_mid_bytes()), apart from the
    // plan's own scratch, which other executions may use between units.
    int progressive_prepare(int want);
    size_t progressive_mid_bytes() const;
    void progressive_unit(int u, void* data, void* mid, bool inverse, hipStream_t s);
    struct impl;
    std::unique_ptr<impl> p;
};

// Real DFT (dft_plan_real): n real values <-> n/2+1 (CCs) or n/2 (Perm) complex values.
template <typename T>
class gpu_dft_real
{
public:
    gpu_dft_real(size_t n, int fmt);
    ~gpu_dft_real();
    size_t size() const;
    size_t complex_size() const;
    int format() const;
    const schedule_info& schedule() const;
    void forward(void* cout, const void* rin, long long batch, hipStream_t s);
    void inverse(void* rout, const void* cin, long long batch, hipStream_t s);
    // Progressive execution of one transform: the half-length complex units plus one unit that repacks
    // (forward, the last unit) or unpacks (inverse, the first unit) the spectrum. z holds n/2 complex
    // values: the input reals (forward) or the output reals (inverse); c holds complex_size() values:
    // the output spectrum (forward) or the input spectrum (inverse).
    int progressive_prepare(int want);
    size_t progressive_mid_bytes() const;
    void progressive_unit(int u, void* z, void* c, void* mid, bool inverse, hipStream_t s);
    struct impl;
    std::unique_ptr<impl> p;
};

// DCT-II / DCT-III (dct_plan), unscaled, real in and out.
template <typename T>
class gpu_dct
{
public:
    explicit gpu_dct(size_t n);
    ~gpu_dct();
    size_t size() const;
    const schedule_info& schedule() const;
    void execute(void* out, const void* in, long long batch, bool inverse, hipStream_t s);
    struct impl;
    std::unique_ptr<impl> p;
};

// Batched linear convolution / correlation of real signals (convolve, correlate): y = x * h with
// |y| = nx + nh - 1, computed with a real FFT of next_pow2(nx + nh - 1) points, plus the adjoint.
template <typename T>
class gpu_conv
{
public:
    gpu_conv(size_t nx, size_t nh);
    ~gpu_conv();
    size_t fft_size() const;
    size_t out_size() const;
    // y[b] = x[b] conv h[b] (h_dist = 0 shares one filter); correlate: h reversed
    void forward(void* y, const void* x, const void* h, long long h_dist, long long batch, bool correlate,
                 hipStream_t s);
    // adjoint of forward (with the same correlate flag): for convolution dx = g corr h (first nx) and
    // dh = g corr x (first nh); for correlation the filter is h reversed, so dx = g conv h and dh is
    // g corr x reversed. The spectrum of x is recomputed, not stored. dh is accumulated over the batch
    // when h_dist == 0.
    void backward(void* dx, void* dh, const void* g, const void* x, const void* h, long long h_dist,
                  long long batch, bool correlate, hipStream_t s);
    struct impl;
    std::unique_ptr<impl> p;
};

// Complex convolution (convolve of complex<T> vectors): complex FFT of S = next_pow2(na+nb-1) points
// with the plan f (of S points) through the device buffers A and B (S values each). keep_imag = false
// reproduces the reference, which returns only the real part.
template <typename T>
void gpu_convolve_complex(void* y, const void* a, size_t na, const void* b, size_t nb, bool correlate,
                          bool keep_imag, gpu_dft<T>& f, void* A, void* B, hipStream_t s);

// Streaming convolution filter of complex signals with complex taps (convolve_filter<complex<T>>),
// state on the device: overlap-save segments through complex FFTs.
template <typename T>
class gpu_conv_filter_complex
{
public:
    gpu_conv_filter_complex(const void* taps_host, size_t ntaps, size_t block_size);
    ~gpu_conv_filter_complex();
    size_t block_size() const;
    void set_data(const void* taps_host, size_t ntaps);
    void reset();
    // process n complex samples (device pointers), any n; keeps state across calls
    void process(void* out, const void* in, size_t n, hipStream_t s);
    struct impl;
    std::unique_ptr<impl> p;
};

// Streaming convolution filter (convolve_filter<T>), state on the device. Real T (T taps, T signal).
template <typename T>
class gpu_conv_filter
{
public:
    gpu_conv_filter(const T* taps_host, size_t ntaps, size_t block_size);
    ~gpu_conv_filter();
    size_t block_size() const;
    void set_data(const T* taps_host, size_t ntaps);
    void reset();
    // process n samples (device pointers), any n; keeps state across calls
    void process(T* out, const T* in, size_t n, hipStream_t s);
    struct impl;
    std::unique_ptr<impl> p;
};

// Elementwise helpers on device data used by the bindings.
template <typename T>
void gpu_scale(void* data, size_t count, T factor, hipStream_t s);

// The library's stream for the synchronous host-pointer entry points (default priority).
hipStream_t library_stream();

} // namespace kfr_amd
