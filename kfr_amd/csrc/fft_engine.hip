// gfx950 DFT engine: complex, real and DCT plans, Bluestein, FFT convolution and its adjoint.
#include "fft_core.hpp"
#include "plan_desc.hpp"

#include <atomic>
#include <cstdio>

namespace kfr_amd
{

// ------------------------------------------------------------------------------------------------
// runtime helpers

static std::atomic<int> g_alloc_countdown{ 0 };

void set_alloc_failure_countdown(int n) { g_alloc_countdown.store(n); }

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw hip_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

dev_mem::~dev_mem()
{
    if (ptr)
        (void)hipFree(ptr);
}

void dev_mem::reset(size_t b)
{
    if (ptr)
    {
        hip_check(hipFree(ptr), "hipFree");
        ptr   = nullptr;
        bytes = 0;
    }
    if (b == 0)
        return;
    const int c = g_alloc_countdown.load();
    if (c > 0 && g_alloc_countdown.fetch_sub(1) == 1)
        throw hip_error("HIP error in hipMalloc: out of memory (injected allocation failure)");
    hip_check(hipMalloc(&ptr, b), "hipMalloc");
    bytes = b;
}

hipStream_t library_stream()
{
    static hipStream_t s = []
    {
        hipStream_t st;
        hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
        return st;
    }();
    return s;
}

std::string schedule_info::text() const
{
    std::string s = "gpu(gfx950): " + algorithm;
    auto list     = [](const std::vector<int>& v)
    {
        std::string r;
        for (size_t i = 0; i < v.size(); ++i)
            r += (i ? "x" : "") + std::to_string(v[i]);
        return r;
    };
    if (n1)
        s += " n1=" + std::to_string(n1) + " n2=" + std::to_string(n2);
    if (bluestein_m)
        s += " m=" + std::to_string(bluestein_m);
    s += " radices=" + list(radices);
    if (threads_per_transform)
        s += " threads/transform=" + std::to_string(threads_per_transform) +
             " transforms/block=" + std::to_string(transforms_per_block);
    if (!contiguous.empty())
        s += "; contiguous batches: " + contiguous;
    return s;
}

// ------------------------------------------------------------------------------------------------
// Bluestein (the reference's dft_arblen_stage_impl, src/dft/dft-impl.hpp:199-271):
// X = c . IFFT_M( FFT_M(x . c, zero padded) . conj(H) ), c[m] = exp(-i pi (m^2 mod 2n)/n),
// H = FFT_M(h)/M with h the chirp mirrored into both ends of an M-point buffer. The inverse uses
// conj(c) and H. M = next_pow2(2n - 1) as in the reference.

// fused: both M-point transforms and the spectrum product in one workgroup, data in LDS throughout
template <typename T, bool INV>
__global__ void __launch_bounds__(max_block_pow2) k_bluestein(lds_geom G, const cx<T>* __restrict__ tw, long long batch,
                                                    const cx<T>* __restrict__ H, xload<T> ld, xstore<T> st,
                                                    fastdiv mdiv)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds            = reinterpret_cast<cx<T>*>(smem_raw);
    const long long first = (long long)blockIdx.x * G.tpb;
    const fft_ctx<T> c    = make_ctx<T>(G, lds, first, batch);
    block_load(G, lds, first, batch, ld, mdiv); // ld carries the chirp and the zero padding
    __syncthreads();
    fft_passes<T, false, false>(G, tw, c); // M is a power of two
    for (int i = c.t; i < G.n; i += G.tpt)
        c.at(i) = INV ? cmul(c.at(i), H[i]) : cmulc(c.at(i), H[i]);
    __syncthreads();
    fft_passes<T, true, false>(G, tw, c);
    block_store(G, lds, first, batch, st, mdiv); // st carries the chirp and the truncation
}

// h: the FP64 chirp mirrored into both ends of M points
__global__ void k_bluestein_h(cx<double>* h, const cx<double>* chirp, long long n, long long M)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M)
        return;
    cx<double> v = { 0.0, 0.0 };
    if (i < n)
        v = chirp[i];
    else if (i > M - n)
        v = chirp[M - i];
    h[i] = v;
}

template <typename T>
__global__ void k_convert_scale(cx<T>* out, const cx<double>* in, long long count, double scale)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        out[i] = { (T)(in[i].re * scale), (T)(in[i].im * scale) };
}

template <typename T>
void bluestein_tables(size_t n, size_t M, cx<T>* chirp, cx<T>* H, hipStream_t s)
{
    dev_mem c64(n * sizeof(cx<double>));
    gen_twiddles(c64.as<cx<double>>(), (long long)n, 2 * (long long)n, 0, 0, tw_square, 0, s);
    hipLaunchKernelGGL(k_convert_scale<T>, dim3(nblocks(n, 256)), dim3(256), 0, s, chirp, c64.as<cx<double>>(),
                       (long long)n, 1.0);
    hip_check(hipGetLastError(), "chirp conversion");
    // H = FFT_M(h) / M evaluated in FP64 and rounded once
    dev_mem h64(M * sizeof(cx<double>));
    hipLaunchKernelGGL(k_bluestein_h, dim3(nblocks(M, 256)), dim3(256), 0, s, h64.as<cx<double>>(),
                       c64.as<cx<double>>(), (long long)n, (long long)M);
    hip_check(hipGetLastError(), "bluestein kernel spectrum");
    {
        core_fft<double> f64(M, s);
        f64.run<false>(load_plain<double>(h64.ptr, 0), store_plain<double>(h64.ptr, 0), 1, s);
        hip_check(hipStreamSynchronize(s), "bluestein kernel spectrum");
    }
    hipLaunchKernelGGL(k_convert_scale<T>, dim3(nblocks(M, 256)), dim3(256), 0, s, H, h64.as<cx<double>>(),
                       (long long)M, 1.0 / (double)M);
    hip_check(hipGetLastError(), "bluestein kernel spectrum");
    hip_check(hipStreamSynchronize(s), "bluestein tables"); // before the FP64 buffers are released
}
template void bluestein_tables<float>(size_t, size_t, cx<float>*, cx<float>*, hipStream_t);
template void bluestein_tables<double>(size_t, size_t, cx<double>*, cx<double>*, hipStream_t);

// element passes of the chained Bluestein over plain M-point buffers, so that both transforms run on
// the power-of-two kernels: y[b][e] = ld.load(b, e) for e < m (the load's chirp, zero padding and
// spectrum products apply), and st.store(b, e, y[b][e]) for e < n. Items b run over y and z.
template <typename T>
__global__ void k_bs_gather(cx<T>* y, long long ydist, xload<T> ld, long long m, long long items)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long b = (long long)blockIdx.y + (long long)gridDim.y * blockIdx.z;
    if (b >= items || e >= m)
        return;
    y[b * ydist + e] = ld.load(b, e);
}

template <typename T>
__global__ void k_bs_scatter(xstore<T> st, const cx<T>* y, long long ydist, long long n, long long items)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long b = (long long)blockIdx.y + (long long)gridDim.y * blockIdx.z;
    if (b >= items || e >= n)
        return;
    st.store(b, e, y[b * ydist + e]);
}

// ------------------------------------------------------------------------------------------------
// complex plan

template <typename T>
struct gpu_dft<T>::impl
{
    enum algo_t
    {
        CORE,
        BS_FUSED,
        BS_CHAIN
    } algo = CORE;
    size_t n = 0;
    schedule_info info;
    std::unique_ptr<core_fft<T>> core; // length n (CORE) or M (BS_CHAIN)
    size_t M = 0;
    stockham bsS;
    int bs_tpb = 1;
    dev_mem bs_tw, chirp, H, Y;
    long long bs_chunk = 1;

    explicit impl(size_t len)
    {
        if (len == 0)
            throw std::logic_error("DFT size must be positive");
        n                 = len;
        hipStream_t s     = library_stream();
        const plan_desc d = describe_dft(sizeof(T) == 4 ? dtype_t::f32 : dtype_t::f64, n, false);
        if (!d.arblen)
        {
            core.reset(new core_fft<T>(n, s));
            info.algorithm = core->fourstep ? "four-step" : "lds";
            info.radices   = core->radices();
            info.n1        = core->n1;
            info.n2        = core->n2;
            if (!core->fourstep)
            {
                info.threads_per_transform = core->S.tpt;
                info.transforms_per_block  = core->tpb;
            }
            if (core->pow2)
                info.contiguous = core->pow2->describe();
            else if (core->mixed)
                info.contiguous = core->mixed->describe();
        }
        else
            init_bluestein(s);
        hip_check(hipStreamSynchronize(s), "plan initialisation");
    }

    void init_bluestein(hipStream_t s)
    {
        M                = next_pow2(2 * n - 1);
        info.bluestein_m = M;
        chirp.reset(n * sizeof(cx<T>));
        H.reset(M * sizeof(cx<T>));
        bluestein_tables<T>(n, M, chirp.as<cx<T>>(), H.as<cx<T>>(), s);
        if (core_fft<T>::fits_lds(M))
        {
            algo = BS_FUSED;
            make_stockham((long long)M, bsS, pmax_of<T>);
            check_pass_radices<T>(bsS, false);
            bs_tpb = core_fft<T>::pick_tpb(bsS);
            bs_tw.reset((size_t)std::max<long long>(1, bsS.tw_count) * sizeof(cx<T>));
            gen_stockham_twiddles(bsS, bs_tw.as<cx<T>>(), s);
            info.algorithm             = "bluestein-fused";
            info.radices               = bsS.radix;
            info.threads_per_transform = bsS.tpt;
            info.transforms_per_block  = bs_tpb;
        }
        else
        {
            algo = BS_CHAIN;
            core.reset(new core_fft<T>(M, s));
            info.algorithm = "bluestein";
            info.radices   = core->radices();
            info.n1        = core->n1;
            info.n2        = core->n2;
            size_t cap = (size_t)256 << 20;
            if (const char* e = std::getenv("KFR_AMD_BLUESTEIN_CHUNK_MB"))
                cap = (size_t)std::max(1, atoi(e)) << 20;
            bs_chunk = (long long)std::max<size_t>(1, cap / (M * sizeof(cx<T>)));
        }
        hip_check(hipStreamSynchronize(s), "bluestein initialisation");
    }

    template <bool INV>
    void run_t(const xload<T>& ld, const xstore<T>& st, long long batch, hipStream_t s, workspace w = {})
    {
        if (batch <= 0)
            return;
        if (algo == CORE)
        {
            core->template run<INV>(ld, st, batch, s, w);
            return;
        }
        if (w.p && algo == BS_CHAIN)
            throw std::logic_error("this Bluestein length keeps intermediates of its own and cannot run in a workspace");
        // the chirp post-multiplies the input and pre-multiplies the output (conjugated for the inverse)
        xload<T> lc  = ld;
        lc.mtab      = chirp.as<cx<T>>();
        lc.mconj     = INV;
        lc.lim       = (long long)n;
        xstore<T> sc = st;
        sc.mtab      = chirp.as<cx<T>>();
        sc.mconj     = INV;
        sc.lim       = (long long)n;
        if (algo == BS_FUSED)
        {
            lds_geom G      = make_geom(bsS, bs_tpb, 0, 1, (int)M);
            const size_t sh = lds_padded_elems((size_t)bs_tpb * M) * sizeof(cx<T>);
            allow_lds(k_bluestein<T, INV>, sh);
            timing::span ts([] { return timing::kernel_name("k_bluestein", timing::type_name<T>(), INV); },
                            2 * (long long)sizeof(cx<T>) * batch * (long long)n, s);
            hipLaunchKernelGGL((k_bluestein<T, INV>), dim3(nblocks(batch, bs_tpb)), dim3(bsS.tpt * bs_tpb), sh, s, G,
                               bs_tw.as<cx<T>>(), batch, H.as<cx<T>>(), lc, sc, make_fastdiv((uint32_t)M));
            hip_check(hipGetLastError(), "bluestein launch");
            return;
        }
        Y.ensure((size_t)std::min(batch, bs_chunk) * M * sizeof(cx<T>));
        for (long long b0 = 0; b0 < batch; b0 += bs_chunk)
        {
            const long long nb = std::min(bs_chunk, batch - b0);
            xload<T> l1        = lc;
            l1.b0 += b0;
            xload<T> l2  = load_plain<T>(Y.ptr, (long long)M);
            l2.mode      = ld_mul;
            l2.q         = H.as<cx<T>>();
            l2.flag      = INV ? 0 : 1;
            xstore<T> s2 = sc;
            s2.b0 += b0;
            if (core->pow2)
            {
                // element passes over plain buffers and both M-point transforms on the power-of-two
                // kernels (measured 2.3x faster for 96001 and 100003 points than the generic kernels
                // doing the element operations in their load and store phases)
                cx<T>* y          = Y.as<cx<T>>();
                const dim3 block(256);
                const dim3 gm(nblocks((long long)M, 256), (unsigned)std::min<long long>(nb, 65535),
                              (unsigned)((nb + 65534) / 65535));
                const dim3 gn(nblocks((long long)n, 256), gm.y, gm.z);
                const long long mb = (long long)sizeof(cx<T>) * nb * (long long)M;
                {
                    timing::span ts([] { return timing::kernel_name("k_bs_gather", timing::type_name<T>()); },
                                    mb + (long long)sizeof(cx<T>) * nb * (long long)n, s);
                    hipLaunchKernelGGL(k_bs_gather<T>, gm, block, 0, s, y, (long long)M, l1, (long long)M, nb);
                    hip_check(hipGetLastError(), "bluestein input launch");
                }
                core->pow2->run(false, y, y, nb, s);
                {
                    timing::span ts([] { return timing::kernel_name("k_bs_gather", timing::type_name<T>()); },
                                    2 * mb, s);
                    hipLaunchKernelGGL(k_bs_gather<T>, gm, block, 0, s, y, (long long)M, l2, (long long)M, nb);
                    hip_check(hipGetLastError(), "bluestein spectrum launch");
                }
                core->pow2->run(true, y, y, nb, s);
                {
                    timing::span ts([] { return timing::kernel_name("k_bs_scatter", timing::type_name<T>()); },
                                    2 * (long long)sizeof(cx<T>) * nb * (long long)n, s);
                    hipLaunchKernelGGL(k_bs_scatter<T>, gn, block, 0, s, s2, y, (long long)M, (long long)n, nb);
                    hip_check(hipGetLastError(), "bluestein output launch");
                }
                continue;
            }
            core->template run<false>(l1, store_plain<T>(Y.ptr, (long long)M), nb, s);
            core->template run<true>(l2, s2, nb, s);
        }
    }

    void run(const xload<T>& ld, const xstore<T>& st, long long batch, bool inverse, hipStream_t s,
             workspace w = {})
    {
        if (inverse)
            run_t<true>(ld, st, batch, s, w);
        else
            run_t<false>(ld, st, batch, s, w);
    }

    // progressive execution (see gpu_dft::progressive_prepare)
    int prog_prepare(int want)
    {
        if (algo != CORE)
            return 1; // Bluestein: the whole transform in one unit
        hipStream_t s = library_stream();
        core->progressive_prepare(want, s);
        hip_check(hipStreamSynchronize(s), "progressive plan");
        return core->progressive_units();
    }
    int prog_units() const { return algo == CORE ? core->progressive_units() : 1; }
    size_t prog_mid_bytes() const { return algo == CORE && core->fourstep ? n * sizeof(cx<T>) : 0; }
    void prog_unit(int u, void* data, void* mid, bool inverse, hipStream_t s)
    {
        if (algo != CORE)
        {
            run(load_plain<T>(data, (long long)n), store_plain<T>(data, (long long)n), 1, inverse, s);
            return;
        }
        if (inverse)
            core->template progressive_unit<true>(u, static_cast<cx<T>*>(data), static_cast<cx<T>*>(mid), s);
        else
            core->template progressive_unit<false>(u, static_cast<cx<T>*>(data), static_cast<cx<T>*>(mid), s);
    }
};

template <typename T>
gpu_dft<T>::gpu_dft(size_t n) : p(new impl(n))
{
}
template <typename T>
gpu_dft<T>::~gpu_dft() = default;
template <typename T>
size_t gpu_dft<T>::size() const
{
    return p->n;
}
template <typename T>
const schedule_info& gpu_dft<T>::schedule() const
{
    return p->info;
}

template <typename T>
void gpu_dft<T>::execute_dist(void* out, long long odist, const void* in, long long idist, long long batch,
                              bool inverse, hipStream_t s)
{
    // In place is safe: the LDS kernel loads a block's transforms before storing them, the four-step
    // column pass writes only the intermediate, and Bluestein stages through its own buffer.
    p->run(load_plain<T>(in, idist), store_plain<T>(out, odist), batch, inverse, s);
}

template <typename T>
void gpu_dft<T>::execute(void* out, const void* in, long long batch, bool inverse, hipStream_t s)
{
    execute_dist(out, (long long)p->n, in, (long long)p->n, batch, inverse, s);
}

template <typename T>
void gpu_dft<T>::execute_workspace(void* out, const void* in, long long batch, bool inverse, hipStream_t s,
                                  workspace w)
{
    p->run(load_plain<T>(in, (long long)p->n), store_plain<T>(out, (long long)p->n), batch, inverse, s, w);
}

template <typename T>
int gpu_dft<T>::progressive_prepare(int want)
{
    return p->prog_prepare(want);
}
template <typename T>
size_t gpu_dft<T>::progressive_mid_bytes() const
{
    return p->prog_mid_bytes();
}
template <typename T>
void gpu_dft<T>::progressive_unit(int u, void* data, void* mid, bool inverse, hipStream_t s)
{
    p->prog_unit(u, data, mid, inverse, s);
}

// ------------------------------------------------------------------------------------------------
// real plan (dft_plan_real): the n real values are read as n/2 complex values z, Z = FFT(z), and the
// spectrum is repacked (to_fmt / from_fmt, src/dft/fft-impl.hpp:1799-1896) with the twiddles
// tw[k] = exp(-i (2 pi k/n + pi/2)) = W_{4n}^{4k+n}.

// NT: nontemporal stores (the repack pass over spectra in global memory: R2C 2^24 x 16 and
// 30,720,000 x 8 measured 1.5-2.7% faster with them)
template <typename T, bool NT = false>
__device__ __forceinline__ void repack_pair(cx<T>* o, const cx<T>* __restrict__ rtw, long long csize, long long k,
                                            cx<T> zk, cx<T> znk)
{
    // to_fmt (src/dft/fft-impl.hpp:1799-1841) for output pair (k, csize-k), 0 < k < (csize+1)/2
    const cx<T> fpnk = cconj(znk);
    const cx<T> f1   = cadd(zk, fpnk);
    const cx<T> f2   = csub(zk, fpnk);
    const cx<T> t    = cmul(f2, rtw[k]);
    const cx<T> a    = cscale(cadd(f1, t), T(0.5));
    const cx<T> b    = cconj(cscale(csub(f1, t), T(0.5)));
    if constexpr (NT)
    {
        using V2 = T __attribute__((ext_vector_type(2)));
        __builtin_nontemporal_store(V2{ a.re, a.im }, reinterpret_cast<V2*>(o + k));
        __builtin_nontemporal_store(V2{ b.re, b.im }, reinterpret_cast<V2*>(o + (csize - k)));
    }
    else
    {
        o[k]         = a;
        o[csize - k] = b;
    }
}
template <typename T>
__device__ __forceinline__ void repack_dc(cx<T>* o, long long csize, int fmt, cx<T> dc)
{
    if (fmt == 1)
    {
        o[0]     = { dc.re + dc.im, T(0) };
        o[csize] = { dc.re - dc.im, T(0) };
    }
    else
        o[0] = { dc.re + dc.im, dc.re - dc.im };
}

// repack of a computed half-length spectrum Z in global memory: one thread per pair
template <typename T>
__global__ void k_real_repack(cx<T>* out, long long odist, const cx<T>* Z, long long zdist,
                              const cx<T>* __restrict__ rtw, long long csize, int fmt, long long batch)
{
    const long long count = (csize + 1) / 2;
    const long long k     = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long b     = (long long)blockIdx.y + (long long)gridDim.y * blockIdx.z;
    if (b >= batch || k > count)
        return;
    const cx<T>* z = Z + b * zdist;
    cx<T>* o       = out + b * odist;
    if (k == 0)
        repack_dc(o, csize, fmt, z[0]);
    else if (k == count)
    {
        if (csize % 2 == 0)
            o[csize / 2] = cconj(z[csize / 2]);
    }
    else
        repack_pair<T, true>(o, rtw, csize, k, z[k], z[csize - k]);
}

// fused forward real transform for LDS-resident half lengths: the FFT result is repacked from LDS
template <typename T, bool MIXED>
__global__ void __launch_bounds__(max_block_of(MIXED)) k_real_fwd(lds_geom G, const cx<T>* __restrict__ tw,
                                                                   long long batch, xload<T> ld, cx<T>* out,
                                                                   long long odist, const cx<T>* __restrict__ rtw,
                                                                   int fmt, fastdiv ndiv)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds            = reinterpret_cast<cx<T>*>(smem_raw);
    const long long first = (long long)blockIdx.x * G.tpb;
    const fft_ctx<T> c    = make_ctx<T>(G, lds, first, batch);
    block_load(G, lds, first, batch, ld, ndiv);
    __syncthreads();
    fft_passes<T, false, MIXED>(G, tw, c);
    if (!c.valid)
        return;
    const long long csize = G.n;
    const long long count = (csize + 1) / 2;
    cx<T>* o              = out + c.g * odist;
    for (long long k = c.t; k <= count; k += G.tpt)
    {
        if (k == 0)
            repack_dc(o, csize, fmt, c.at(0));
        else if (k == count)
        {
            if (csize % 2 == 0)
                o[csize / 2] = cconj(c.at((int)(csize / 2)));
        }
        else
            repack_pair(o, rtw, csize, k, c.at((int)k), c.at((int)(csize - k)));
    }
}

template <typename T>
struct gpu_dft_real<T>::impl
{
    size_t n = 0, csize = 0;
    int fmt = 1;
    std::unique_ptr<typename gpu_dft<T>::impl> c; // complex n/2 plan
    dev_mem rtw;
    dev_mem Z;

    impl(size_t len, int f)
    {
        if (len % 2 != 0 || len == 0)
            throw std::logic_error("dft_plan_real requires size to be even");
        n     = len;
        csize = len / 2;
        fmt   = f;
        c.reset(new typename gpu_dft<T>::impl(csize));
        hipStream_t s         = library_stream();
        const long long count = (long long)(csize + 1) / 2 + 1;
        rtw.reset((size_t)count * sizeof(cx<T>));
        gen_twiddles(rtw.as<cx<T>>(), count, 4 * (long long)n, 4, (long long)n, tw_linear, 0, s);
        hip_check(hipStreamSynchronize(s), "real plan initialisation");
    }
    size_t complex_size() const { return fmt == 1 ? csize + 1 : csize; }

    bool fused_forward() const { return c->algo == gpu_dft<T>::impl::CORE && !c->core->fourstep; }

    // the power-of-two plan of the half length, when it has one-kernel real transforms
    pow2_fft<T>* pow2_half() const
    {
        return c->algo == gpu_dft<T>::impl::CORE && c->core && c->core->pow2 ? c->core->pow2.get() : nullptr;
    }
    static bool plain_ld(const xload<T>& l, long long dist)
    {
        return l.mode == ld_plain && l.src == 0 && l.q == nullptr && l.mtab == nullptr && l.lim == LLONG_MAX &&
               l.pdist == dist;
    }

    void forward_op(const xload<T>& ld, cx<T>* out, long long odist, long long batch, hipStream_t s)
    {
        if (batch <= 0)
            return;
        // power-of-two halves in one workgroup: transform and repack in one k_pow2-style kernel
        if (pow2_fft<T>* h = pow2_half(); h && plain_ld(ld, (long long)csize) &&
            h->real(false, static_cast<const cx<T>*>(ld.p) + ld.b0 * (long long)csize, (long long)csize, out, odist,
                    rtw.ptr, fmt, batch, s))
            return;
        if (fused_forward())
        {
            core_fft<T>& f  = *c->core;
            lds_geom G      = make_geom(f.S, f.tpb, 0, 1, (int)csize);
            const size_t sh = lds_padded_elems((size_t)f.tpb * csize) * sizeof(cx<T>);
            const dim3 grid(nblocks(batch, f.tpb)), block(f.S.tpt * f.tpb);
            const fastdiv nd = make_fastdiv((uint32_t)csize);
            const bool mx    = is_mixed(f.S);
            timing::span ts([mx] { return timing::kernel_name("k_real_fwd", timing::type_name<T>(), mx); },
                            2 * (long long)sizeof(cx<T>) * batch * (long long)csize, s);
            if (mx)
            {
                allow_lds(k_real_fwd<T, true>, sh);
                hipLaunchKernelGGL((k_real_fwd<T, true>), grid, block, sh, s, G, f.tw.template as<cx<T>>(), batch, ld,
                                   out, odist, rtw.template as<cx<T>>(), fmt, nd);
            }
            else
            {
                allow_lds(k_real_fwd<T, false>, sh);
                hipLaunchKernelGGL((k_real_fwd<T, false>), grid, block, sh, s, G, f.tw.template as<cx<T>>(), batch, ld,
                                   out, odist, rtw.template as<cx<T>>(), fmt, nd);
            }
            hip_check(hipGetLastError(), "real forward launch");
            return;
        }
        // rounds of at most 1 GiB of half-length spectra, as the complex plans inside (measured: R2C
        // 2^24 x 16 12% and 30,720,000 x 8 4.5% faster than 256 MiB rounds, whose extra launches each
        // pay a ramp-up and drain)
        const long long per = (long long)csize;
        long long cap       = (long long)1 << 30;
        if (const char* e = std::getenv("KFR_AMD_REAL_CHUNK_MB"))
            cap = (long long)std::max(1, atoi(e)) << 20;
        const long long chunk = std::max<long long>(1, cap / (per * (long long)sizeof(cx<T>)));
        Z.ensure((size_t)std::min(chunk, batch) * per * sizeof(cx<T>));
        for (long long b0 = 0; b0 < batch; b0 += chunk)
        {
            const long long nb = std::min(chunk, batch - b0);
            xload<T> l         = ld;
            l.b0 += b0;
            c->run(l, store_plain<T>(Z.ptr, per), nb, false, s);
            const long long count = (long long)(csize + 1) / 2;
            dim3 grid(nblocks(count + 1, 256), (unsigned)std::min<long long>(nb, 65535),
                      (unsigned)((nb + 65534) / 65535));
            timing::span ts([] { return timing::kernel_name("k_real_repack", timing::type_name<T>()); },
                            2 * (long long)sizeof(cx<T>) * nb * per, s);
            hipLaunchKernelGGL(k_real_repack<T>, grid, dim3(256), 0, s, out + b0 * odist, odist, Z.as<cx<T>>(), per,
                               rtw.as<cx<T>>(), (long long)csize, fmt, nb);
            hip_check(hipGetLastError(), "real repack launch");
        }
    }

    // inverse: `src` describes the half spectrum (plain or a product of two spectra)
    void inverse_op(xload<T> src, const xstore<T>& st, long long batch, hipStream_t s)
    {
        if (batch <= 0)
            return;
        if (pow2_fft<T>* h = pow2_half(); h && plain_ld(src, (long long)complex_size()) && st.mode == st_plain &&
            st.pdist == (long long)csize && st.mtab == nullptr && st.lim == LLONG_MAX &&
            h->real(true, static_cast<const cx<T>*>(src.p) + src.b0 * src.pdist, src.pdist,
                    static_cast<cx<T>*>(st.p) + st.b0 * (long long)csize, (long long)csize, rtw.ptr, fmt, batch, s))
            return;
        src.mode = ld_c2r;
        src.r    = rtw.as<cx<T>>();
        src.n    = (long long)csize;
        src.flag = fmt;
        if (fused_forward())
        {
            // one workgroup per transform: the unpacking runs in the load phase of the same kernel
            c->run(src, st, batch, true, s);
            return;
        }
        // Longer lengths: the from_fmt unpacking as a pass of its own into plain half-length
        // sequences, then the complex plan's plain path (power-of-two, menu and multi-pass kernels),
        // in rounds of at most 1 GiB of half-length data like the forward. With the unpacking in the
        // load phase these lengths ran the generic four-step instead: C2R took 3.0x the R2C time at
        // 2^24 x 16 and 3.4x at 30,720,000 x 8.
        const long long per = (long long)csize;
        long long cap       = (long long)1 << 30;
        if (const char* e = std::getenv("KFR_AMD_REAL_CHUNK_MB"))
            cap = (long long)std::max(1, atoi(e)) << 20;
        const long long chunk = std::max<long long>(1, cap / (per * (long long)sizeof(cx<T>)));
        Z.ensure((size_t)std::min(chunk, batch) * per * sizeof(cx<T>));
        for (long long b0 = 0; b0 < batch; b0 += chunk)
        {
            const long long nb = std::min(chunk, batch - b0);
            xload<T> l         = src;
            l.b0 += b0;
            const dim3 grid(nblocks(per, 256), (unsigned)std::min<long long>(nb, 65535),
                            (unsigned)((nb + 65534) / 65535));
            {
                timing::span ts([] { return timing::kernel_name("k_bs_gather", timing::type_name<T>()); },
                                (long long)sizeof(cx<T>) * nb * ((long long)complex_size() + per), s);
                hipLaunchKernelGGL(k_bs_gather<T>, grid, dim3(256), 0, s, Z.as<cx<T>>(), per, l, per, nb);
            }
            hip_check(hipGetLastError(), "real unpack launch");
            xstore<T> o = st;
            o.b0 += b0;
            c->run(load_plain<T>(Z.ptr, per), o, nb, true, s);
        }
    }
};

template <typename T>
gpu_dft_real<T>::gpu_dft_real(size_t n, int fmt) : p(new impl(n, fmt))
{
}
template <typename T>
gpu_dft_real<T>::~gpu_dft_real() = default;
template <typename T>
size_t gpu_dft_real<T>::size() const
{
    return p->n;
}
template <typename T>
size_t gpu_dft_real<T>::complex_size() const
{
    return p->complex_size();
}
template <typename T>
int gpu_dft_real<T>::format() const
{
    return p->fmt;
}
template <typename T>
const schedule_info& gpu_dft_real<T>::schedule() const
{
    return p->c->info;
}
template <typename T>
void gpu_dft_real<T>::forward(void* cout, const void* rin, long long batch, hipStream_t s)
{
    p->forward_op(load_plain<T>(rin, (long long)p->csize), static_cast<cx<T>*>(cout), (long long)p->complex_size(),
                  batch, s);
}
template <typename T>
void gpu_dft_real<T>::inverse(void* rout, const void* cin, long long batch, hipStream_t s)
{
    xload<T> src = load_plain<T>(cin, (long long)p->complex_size());
    p->inverse_op(src, store_plain<T>(rout, (long long)p->csize), batch, s);
}

template <typename T>
int gpu_dft_real<T>::progressive_prepare(int want)
{
    return p->c->prog_prepare(std::max(1, want - 1)) + 1;
}
template <typename T>
size_t gpu_dft_real<T>::progressive_mid_bytes() const
{
    return p->c->prog_mid_bytes();
}
template <typename T>
void gpu_dft_real<T>::progressive_unit(int u, void* z, void* c, void* mid, bool inverse, hipStream_t s)
{
    const int half      = p->c->prog_units();
    const long long cs  = (long long)p->csize;
    if (!inverse)
    {
        if (u < half)
        {
            p->c->prog_unit(u, z, mid, false, s);
            return;
        }
        // to_fmt repack of the half spectrum (src/dft/fft-impl.hpp:1799-1841)
        const long long count = (cs + 1) / 2;
        hipLaunchKernelGGL(k_real_repack<T>, dim3(nblocks(count + 1, 256)), dim3(256), 0, s, static_cast<cx<T>*>(c),
                           (long long)p->complex_size(), static_cast<const cx<T>*>(z), cs, p->rtw.template as<cx<T>>(),
                           cs, p->fmt, 1ll);
        hip_check(hipGetLastError(), "real repack launch");
        return;
    }
    if (u == 0)
    {
        // from_fmt unpacking (src/dft/fft-impl.hpp:1853-1896) into the half-length sequence
        xload<T> src = load_plain<T>(c, (long long)p->complex_size());
        src.mode     = ld_c2r;
        src.r        = p->rtw.template as<cx<T>>();
        src.n        = cs;
        src.flag     = p->fmt;
        hipLaunchKernelGGL(k_bs_gather<T>, dim3(nblocks(cs, 256)), dim3(256), 0, s, static_cast<cx<T>*>(z), cs, src,
                           cs, 1ll);
        hip_check(hipGetLastError(), "real unpack launch");
        return;
    }
    p->c->prog_unit(u - 1, z, mid, true, s);
}

// ------------------------------------------------------------------------------------------------
// DCT (dct_plan, include/kfr/dft/fft.hpp:927-970): DCT-II through an n-point complex DFT of the
// even/odd mirrored input, y_k = Re(Y_k W_{4n}^k); the inverse (DCT-III) pre-multiplies by
// W_{4n}^k with x_0 halved and de-interleaves the real parts.

// Lengths n = 4m whose half length has a plain fast path use a real transform of the permuted input
// instead (Makhoul): v = x mirrored (v[e] = x[2e], v[n-1-e] = x[2e+1]) is packed as n/2 complex values
// z[q] = (v[2q], v[2q+1]) by one coalesced pass, transformed with the n/2-point plan, and one pass
// forms the real spectrum V[k] = A[k] + W_n^k B[k] (A, B the spectra of the even and odd samples,
// from Z[k] and Z[n/2-k]) and writes y[k] = Re(V[k] W_{4n}^k) and y[n-k] = -Im(V[k] W_{4n}^k). DCT-III
// runs the same steps backwards: the spectrum of the symmetrised pre-twiddled input, an n/2-point
// inverse transform, and the permutation back.

// z[q] = (x[4q], x[4q+2]), z[n/2-1-q] = (x[4q+3], x[4q+1]) for q < n/4: the mirror permutation
template <typename T>
__global__ void k_dct2_pack(cx<T>* z, const T* x, long long n, long long count)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long quarter = n / 4;
    if (i >= count * quarter)
        return;
    const long long b = i / quarter;
    const long long q = i - b * quarter;
    const T* xb       = x + b * n;
    cx<T>* zb         = z + b * (n / 2);
    const T x0 = xb[4 * q], x1 = xb[4 * q + 1], x2 = xb[4 * q + 2], x3 = xb[4 * q + 3];
    zb[q]              = { x0, x2 };
    zb[n / 2 - 1 - q]  = { x3, x1 };
}

// inverse of k_dct2_pack: x[4q..4q+3] = (u[q].re, u[h-1-q].im, u[q].im, u[h-1-q].re)
template <typename T>
__global__ void k_dct3_unpack(T* x, const cx<T>* u, long long n, long long count)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long quarter = n / 4;
    if (i >= count * quarter)
        return;
    const long long b = i / quarter;
    const long long q = i - b * quarter;
    T* xb             = x + b * n;
    const cx<T>* ub   = u + b * (n / 2);
    const cx<T> a = ub[q], c = ub[n / 2 - 1 - q];
    xb[4 * q]     = a.re;
    xb[4 * q + 1] = c.im;
    xb[4 * q + 2] = a.im;
    xb[4 * q + 3] = c.re;
}

// DCT-II output from Z = FFT_{n/2}(z): thread k (k <= n/4) handles the pair k, n/2 - k
template <typename T>
__global__ void k_dct2_post(T* y, const cx<T>* Z, const cx<T>* wn, const cx<T>* w4n, long long n, long long count)
{
    const long long i  = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long h  = n / 2;
    const long long pr = n / 4 + 1;
    if (i >= count * pr)
        return;
    const long long b = i / pr;
    const long long k = i - b * pr;
    const cx<T>* Zb   = Z + b * h;
    T* yb             = y + b * n;
    const cx<T> zk = Zb[k == 0 ? 0 : k], zm = Zb[k == 0 ? 0 : h - k];
    // V[k] = (Zk + conj Zm)/2 + W_n^k (Zk - conj Zm)/(2i); V[h-k] likewise with the roles swapped
    auto spec = [&](cx<T> a, cx<T> c, long long kk) -> cx<T>
    {
        const cx<T> A  = { T(0.5) * (a.re + c.re), T(0.5) * (a.im - c.im) };
        const cx<T> d  = { a.re - c.re, a.im + c.im }; // a - conj c
        const cx<T> B  = { T(0.5) * d.im, -T(0.5) * d.re }; // d / (2i)
        const cx<T> w  = wn[kk];
        return { A.re + w.re * B.re - w.im * B.im, A.im + w.re * B.im + w.im * B.re };
    };
    auto emit = [&](cx<T> V, long long kk)
    {
        const cx<T> w = w4n[kk];
        const T re    = V.re * w.re - V.im * w.im;
        const T im    = V.re * w.im + V.im * w.re;
        yb[kk] = re;
        if (kk > 0 && n - kk != kk)
            yb[n - kk] = -im;
    };
    if (k == 0)
    {
        // V[0] = Re Z0 + Im Z0 (both real), V[h] = Re Z0 - Im Z0
        emit({ zk.re + zk.im, T(0) }, 0);
        const cx<T> Vh = { zk.re - zk.im, T(0) };
        emit(Vh, h);
        return;
    }
    emit(spec(zk, zm, k), k);
    if (h - k != k)
        emit(spec(zm, zk, h - k), h - k);
}

// DCT-III input side: V[k] = y'[k] W_{4n}^k (y'[0] = y[0]/2); the real part of DFT_n(V) is the
// unscaled inverse real transform of H[k] = (conj V[k] + V[n-k])/2, whose n/2-point packing is
// Z'[k] = (H[k] + conj H[h-k]) + i (H[k] - conj H[h-k]) conj(W_n^k)
template <typename T>
__global__ void k_dct3_pre(cx<T>* Zp, const T* y, const cx<T>* wn, const cx<T>* w4n, long long n, long long count)
{
    const long long i  = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long h  = n / 2;
    const long long pr = n / 4 + 1;
    if (i >= count * pr)
        return;
    const long long b = i / pr;
    const long long k = i - b * pr;
    const T* yb       = y + b * n;
    cx<T>* Zb         = Zp + b * h;
    auto Hs = [&](long long kk) -> cx<T>
    {
        // H[kk] = (conj V[kk] + V[n-kk]) / 2, V[n] = V[0]
        const T a0        = yb[kk] * (kk == 0 ? T(0.5) : T(1));
        const cx<T> w0    = w4n[kk];
        const cx<T> V0    = { a0 * w0.re, a0 * w0.im };
        const long long m = kk == 0 ? 0 : n - kk;
        const T a1        = yb[m] * (m == 0 ? T(0.5) : T(1));
        const cx<T> w1    = w4n[m];
        const cx<T> V1    = { a1 * w1.re, a1 * w1.im };
        return { T(0.5) * (V0.re + V1.re), T(0.5) * (V1.im - V0.im) };
    };
    auto pack = [&](cx<T> a, cx<T> c, long long kk) -> cx<T>
    {
        // (a + conj c) + i (a - conj c) conj(W_n^kk)
        const cx<T> s = { a.re + c.re, a.im - c.im };
        const cx<T> d = { a.re - c.re, a.im + c.im };
        const cx<T> w = wn[kk];
        const cx<T> e = { d.re * w.re + d.im * w.im, d.im * w.re - d.re * w.im }; // d conj(w)
        return { s.re - e.im, s.im + e.re };
    };
    const cx<T> Hk = Hs(k);
    const cx<T> Hm = Hs(h - k);
    Zb[k == h ? 0 : k] = pack(Hk, Hm, k);
    if (k != 0 && h - k != k)
        Zb[h - k] = pack(Hm, Hk, h - k);
}

template <typename T>
struct gpu_dct<T>::impl
{
    size_t n = 0;
    std::unique_ptr<typename gpu_dft<T>::impl> c;  // n-point complex plan (generic path)
    std::unique_ptr<typename gpu_dft<T>::impl> ch; // n/2-point complex plan (real-transform path)
    dev_mem w;  // W_{4n}^k, k < n
    dev_mem wn; // W_n^k, k <= n/2
    dev_mem z;  // packed n/2-point data of one round
    long long chunk = 1;
    bool fused      = false; // DCT-II and DCT-III in one kernel (power-of-two n/2 within one workgroup)
    explicit impl(size_t len)
    {
        n             = len;
        hipStream_t s = library_stream();
        w.reset(std::max<size_t>(1, n) * sizeof(cx<T>));
        gen_twiddles(w.as<cx<T>>(), (long long)n, 4 * (long long)n, 1, 0, tw_linear, 0, s);
        const size_t h = n / 2;
        if (n % 4 == 0 && n >= 16 && (pow2_fft<T>::supports(h) || mixed_fft<T>::supports(h)))
        {
            ch.reset(new typename gpu_dft<T>::impl(h));
            wn.reset((h + 1) * sizeof(cx<T>));
            gen_twiddles(wn.as<cx<T>>(), (long long)h + 1, (long long)n, 1, 0, tw_linear, 0, s);
            chunk = (long long)std::max<size_t>(1, ((size_t)512 << 20) / (h * sizeof(cx<T>)));
            fused = ch->core && ch->core->pow2 && ch->core->pow2->dct(false, nullptr, nullptr, nullptr, nullptr, 0, s);
        }
        else
            c.reset(new typename gpu_dft<T>::impl(n));
        hip_check(hipStreamSynchronize(s), "dct plan initialisation");
    }
};

template <typename T>
gpu_dct<T>::gpu_dct(size_t n) : p(new impl(n))
{
}
template <typename T>
gpu_dct<T>::~gpu_dct() = default;
template <typename T>
size_t gpu_dct<T>::size() const
{
    return p->n;
}
template <typename T>
const schedule_info& gpu_dct<T>::schedule() const
{
    return p->ch ? p->ch->info : p->c->info;
}
template <typename T>
void gpu_dct<T>::execute(void* out, const void* in, long long batch, bool inverse, hipStream_t s)
{
    const long long n = (long long)p->n;
    if (p->ch)
    {
        impl& q           = *p;
        const long long h = n / 2;
        if (q.fused)
        {
            q.ch->core->pow2->dct(inverse, in, out, q.wn.ptr, q.w.ptr, batch, s);
            return;
        }
        q.z.ensure((size_t)(std::min(batch, q.chunk) * h) * sizeof(cx<T>));
        cx<T>* z = q.z.template as<cx<T>>();
        for (long long b0 = 0; b0 < batch; b0 += q.chunk)
        {
            const long long nb = std::min(q.chunk, batch - b0);
            const T* x         = static_cast<const T*>(in) + b0 * n;
            T* y               = static_cast<T*>(out) + b0 * n;
            const long long pr = n / 4 + 1;
            if (!inverse)
            {
                {
                    timing::span ts([] { return timing::kernel_name("k_dct2_pack", timing::type_name<T>()); },
                                    2 * (long long)sizeof(T) * nb * n, s);
                    hipLaunchKernelGGL(k_dct2_pack<T>, dim3(nblocks(nb * (n / 4), 256)), dim3(256), 0, s, z, x, n,
                                       nb);
                }
                hip_check(hipGetLastError(), "dct pack");
                q.ch->run(load_plain<T>(z, h), store_plain<T>(z, h), nb, false, s);
                {
                    timing::span ts([] { return timing::kernel_name("k_dct2_post", timing::type_name<T>()); },
                                    2 * (long long)sizeof(T) * nb * n, s);
                    hipLaunchKernelGGL(k_dct2_post<T>, dim3(nblocks(nb * pr, 256)), dim3(256), 0, s, y, z,
                                       q.wn.template as<cx<T>>(), q.w.template as<cx<T>>(), n, nb);
                }
                hip_check(hipGetLastError(), "dct post");
            }
            else
            {
                {
                    timing::span ts([] { return timing::kernel_name("k_dct3_pre", timing::type_name<T>()); },
                                    2 * (long long)sizeof(T) * nb * n, s);
                    hipLaunchKernelGGL(k_dct3_pre<T>, dim3(nblocks(nb * pr, 256)), dim3(256), 0, s, z, x,
                                       q.wn.template as<cx<T>>(), q.w.template as<cx<T>>(), n, nb);
                }
                hip_check(hipGetLastError(), "dct-iii pre");
                q.ch->run(load_plain<T>(z, h), store_plain<T>(z, h), nb, true, s);
                {
                    timing::span ts([] { return timing::kernel_name("k_dct3_unpack", timing::type_name<T>()); },
                                    2 * (long long)sizeof(T) * nb * n, s);
                    hipLaunchKernelGGL(k_dct3_unpack<T>, dim3(nblocks(nb * (n / 4), 256)), dim3(256), 0, s, y, z,
                                       n, nb);
                }
                hip_check(hipGetLastError(), "dct-iii unpack");
            }
        }
        return;
    }
    xload<T> l;
    xstore<T> st;
    l.p  = in;
    l.n  = n;
    st.p = out;
    st.n = n;
    if (!inverse)
    {
        l.mode  = ld_dct2;
        st.mode = st_dct2;
        st.q    = p->w.template as<cx<T>>();
    }
    else
    {
        l.mode  = ld_dct3;
        l.q     = p->w.template as<cx<T>>();
        st.mode = st_dct3;
    }
    p->c->run(l, st, batch, false, s);
}

// ------------------------------------------------------------------------------------------------
// FFT convolution of real signals (convolve / correlate, src/dft/convolution-impl.cpp:106-128),
// computed with a real FFT of S = next_pow2(nx + nh - 1) points, and its adjoint.

template <typename T>
inline xload<T> load_real_pad(const void* x, long long dist, long long nvalid, bool reverse)
{
    xload<T> l;
    l.mode  = ld_real_pad;
    l.p     = x;
    l.pdist = dist;
    l.n     = nvalid;
    l.flag  = reverse ? 1 : 0;
    return l;
}
template <typename T>
inline xstore<T> store_real_trunc(void* y, long long dist, long long nvalid, T scale)
{
    xstore<T> s;
    s.mode  = st_real_trunc;
    s.p     = y;
    s.pdist = dist;
    s.n     = nvalid;
    s.scale = scale;
    return s;
}
// half spectrum a*b*scale (conj: a*conj(b)*scale)
template <typename T>
inline xload<T> spec_product(const void* a, long long adist, const cx<T>* b, long long bdist, T scale, bool conj)
{
    xload<T> l;
    l.p     = a;
    l.pdist = adist;
    l.q     = b;
    l.qdist = bdist;
    l.scale = scale;
    l.src   = conj ? 2 : 1;
    return l;
}

template <typename T>
struct gpu_conv<T>::impl
{
    size_t nx = 0, nh = 0, ny = 0, S = 0;
    std::unique_ptr<typename gpu_dft_real<T>::impl> r;
    // short filters against long signals: overlap-save in one kernel per call
    std::unique_ptr<os_conv<T>> os;
    dev_mem X, Hs, G, tmp;
    impl(size_t a, size_t b)
    {
        if (a == 0 || b == 0)
            throw std::logic_error("convolution inputs must not be empty");
        nx = a;
        nh = b;
        ny = a + b - 1;
        S  = std::max<size_t>(2, next_pow2(ny));
        const int L = os_conv<T>::block_log(nx, nh);
        if (L >= 0 && ((size_t)1 << L) < S)
        {
            hipStream_t s = library_stream();
            os.reset(new os_conv<T>(nx, nh, s));
            hip_check(hipStreamSynchronize(s), "convolution plan initialisation");
        }
        else
            r.reset(new typename gpu_dft_real<T>::impl(S, 1));
    }
    long long cs() const { return (long long)(S / 2 + 1); }
};

template <typename T>
gpu_conv<T>::gpu_conv(size_t nx, size_t nh) : p(new impl(nx, nh))
{
}
template <typename T>
gpu_conv<T>::~gpu_conv() = default;
template <typename T>
size_t gpu_conv<T>::fft_size() const
{
    return p->S;
}
template <typename T>
size_t gpu_conv<T>::out_size() const
{
    return p->ny;
}

template <typename T>
void gpu_conv<T>::forward(void* y, const void* x, const void* h, long long h_dist, long long batch, bool correlate,
                          hipStream_t s)
{
    impl& q = *p;
    if (batch <= 0)
        return;
    if (q.os)
    {
        q.os->forward(static_cast<T*>(y), static_cast<const T*>(x), static_cast<const T*>(h), h_dist, batch,
                      correlate, s);
        return;
    }
    const long long cs = q.cs();
    const long long hb = h_dist == 0 ? 1 : batch;
    q.X.ensure((size_t)batch * cs * sizeof(cx<T>));
    q.Hs.ensure((size_t)hb * cs * sizeof(cx<T>));
    q.r->forward_op(load_real_pad<T>(x, (long long)q.nx, (long long)q.nx, false), q.X.template as<cx<T>>(), cs,
                    batch, s);
    q.r->forward_op(load_real_pad<T>(h, h_dist == 0 ? 0 : (long long)q.nh, (long long)q.nh, correlate),
                    q.Hs.template as<cx<T>>(), cs, hb, s);
    q.r->inverse_op(spec_product<T>(q.X.ptr, cs, q.Hs.template as<cx<T>>(), h_dist == 0 ? 0 : cs, T(1) / T(q.S), false),
                    store_real_trunc<T>(y, (long long)q.ny, (long long)q.ny, T(1)), batch, s);
}

template <typename T>
__global__ void k_zero(T* p, long long count);

// a[b*n + j] <-> a[b*n + n-1-j] for every item b < items: thread i swaps pair j = i mod (n/2)
template <typename T>
__global__ void k_reverse_items(T* a, long long n, long long items)
{
    const long long half = n / 2;
    const long long i    = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (half == 0 || i >= items * half)
        return;
    const long long b = i / half;
    const long long j = i - b * half;
    T* p              = a + b * n;
    const T t         = p[j];
    p[j]              = p[n - 1 - j];
    p[n - 1 - j]      = t;
}

template <typename T>
__global__ void k_sum_batch(T* out, const T* in, long long n, long long batch)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    T acc = T(0);
    for (long long b = 0; b < batch; ++b)
        acc += in[b * n + i];
    out[i] = acc;
}

template <typename T>
void gpu_conv<T>::backward(void* dx, void* dh, const void* g, const void* x, const void* h, long long h_dist,
                           long long batch, bool correlate, hipStream_t s)
{
    impl& q = *p;
    if (batch <= 0)
    {
        // no signals: the gradient of a shared filter is the empty sum
        if (dh && h_dist == 0)
        {
            hipLaunchKernelGGL(k_zero<T>, dim3(nblocks((long long)q.nh, 256)), dim3(256), 0, s, static_cast<T*>(dh),
                               (long long)q.nh);
            hip_check(hipGetLastError(), "gradient clear");
        }
        return;
    }
    // the correlation's filter is h reversed: its gradient is the convolution's, reversed
    const long long dh_items = h_dist == 0 ? 1 : batch;
    auto reverse_dh          = [&]
    {
        if (!dh || !correlate)
            return;
        hipLaunchKernelGGL(k_reverse_items<T>, dim3(nblocks(dh_items * (long long)(q.nh / 2), 256) + 1), dim3(256), 0,
                           s, static_cast<T*>(dh), (long long)q.nh, dh_items);
        hip_check(hipGetLastError(), "gradient reversal");
    };
    if (q.os)
    {
        q.os->backward(static_cast<T*>(dx), static_cast<T*>(dh), static_cast<const T*>(g), static_cast<const T*>(x),
                       static_cast<const T*>(h), h_dist, batch, correlate, s);
        reverse_dh();
        return;
    }
    const long long cs = q.cs();
    const long long hb = h_dist == 0 ? 1 : batch;
    q.G.ensure((size_t)batch * cs * sizeof(cx<T>));
    q.Hs.ensure((size_t)hb * cs * sizeof(cx<T>));
    const T scale = T(1) / T(q.S);
    q.r->forward_op(load_real_pad<T>(g, (long long)q.ny, (long long)q.ny, false), q.G.template as<cx<T>>(), cs, batch,
                    s);
    if (dx)
    {
        // dx = g correlated with the filter f (h, or h reversed for the correlation): IFFT(G conj(F)), first
        // nx samples
        q.r->forward_op(load_real_pad<T>(h, h_dist == 0 ? 0 : (long long)q.nh, (long long)q.nh, correlate),
                        q.Hs.template as<cx<T>>(), cs, hb, s);
        q.r->inverse_op(spec_product<T>(q.G.ptr, cs, q.Hs.template as<cx<T>>(), h_dist == 0 ? 0 : cs, scale, true),
                        store_real_trunc<T>(dx, (long long)q.nx, (long long)q.nx, T(1)), batch, s);
    }
    if (dh)
    {
        // dh = g correlated with x: IFFT(G conj(X)), first nh samples; X is recomputed here
        q.X.ensure((size_t)batch * cs * sizeof(cx<T>));
        q.r->forward_op(load_real_pad<T>(x, (long long)q.nx, (long long)q.nx, false), q.X.template as<cx<T>>(), cs,
                        batch, s);
        if (h_dist != 0)
            q.r->inverse_op(spec_product<T>(q.G.ptr, cs, q.X.template as<cx<T>>(), cs, scale, true),
                            store_real_trunc<T>(dh, (long long)q.nh, (long long)q.nh, T(1)), batch, s);
        else
        {
            // shared taps: the gradient sums over the batch
            q.tmp.ensure((size_t)batch * q.nh * sizeof(T));
            q.r->inverse_op(spec_product<T>(q.G.ptr, cs, q.X.template as<cx<T>>(), cs, scale, true),
                            store_real_trunc<T>(q.tmp.ptr, (long long)q.nh, (long long)q.nh, T(1)), batch, s);
            hipLaunchKernelGGL(k_sum_batch<T>, dim3(nblocks((long long)q.nh, 256)), dim3(256), 0, s,
                               static_cast<T*>(dh), q.tmp.template as<T>(), (long long)q.nh, batch);
            hip_check(hipGetLastError(), "gradient reduction");
        }
        reverse_dh();
    }
}

// complex convolution: complex FFT of S points; keep_imag = false returns the real part as the
// reference does (src/dft/convolution-impl.cpp:124-127)
template <typename T>
void gpu_convolve_complex(void* y, const void* a, size_t na, const void* b, size_t nb, bool correlate,
                          bool keep_imag, gpu_dft<T>& plan, void* A, void* B, hipStream_t s)
{
    const size_t ny = na + nb - 1;
    const size_t S  = next_pow2(ny);
    if (plan.size() != S)
        throw std::logic_error("complex convolution: plan size does not match");
    typename gpu_dft<T>::impl& f = *plan.p;
    xload<T> la;
    la.mode = ld_cpad;
    la.p    = a;
    la.n    = (long long)na;
    f.run(la, store_plain<T>(A, 0), 1, false, s);
    xload<T> lb = la;
    lb.p        = b;
    lb.n        = (long long)nb;
    lb.flag     = correlate ? 1 : 0;
    f.run(lb, store_plain<T>(B, 0), 1, false, s);
    xload<T> lm = load_plain<T>(A, 0);
    lm.mode     = ld_mul;
    lm.q        = static_cast<const cx<T>*>(B);
    xstore<T> sy;
    sy.mode  = st_ctrunc;
    sy.p     = y;
    sy.n     = (long long)ny;
    sy.scale = T(1) / T(S);
    sy.flag  = keep_imag ? 1 : 0;
    f.run(lm, sy, 1, true, s);
    hip_check(hipStreamSynchronize(s), "complex convolution");
}

// ------------------------------------------------------------------------------------------------
// Streaming convolution filter (convolve_filter<T>, src/dft/convolution-impl.cpp:37-228). The
// reference's uniformly partitioned overlap-add produces y[t] = sum_k h[k] x[t-k] across calls. Here
// each call is one overlap-save batch: the call's input is appended to the last ntaps-1 samples of
// history, cut into segments of S points overlapping by ntaps-1, and all segments are filtered by
// one batched real FFT, a spectrum product fused into the inverse FFT's load, and an inverse FFT;
// the last S-(ntaps-1) samples of each segment are the output.

template <typename T>
__global__ void k_gather_segments(T* seg, const T* hist, const T* in, long long nh1, long long n, long long S,
                                  long long L, long long nseg)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nseg * S)
        return;
    const long long c = idx / S, j = idx - c * S;
    const long long pos = c * L + j; // position in [history | input]
    T v{};
    if (pos < nh1)
        v = hist[pos];
    else if (pos - nh1 < n)
        v = in[pos - nh1];
    seg[idx] = v;
}

template <typename T>
__global__ void k_scatter_valid(T* out, const T* y, long long nh1, long long n, long long S, long long L,
                                long long nseg)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nseg * L)
        return;
    const long long c = idx / L, j = idx - c * L;
    const long long t = c * L + j;
    if (t < n)
        out[t] = y[c * S + nh1 + j];
}

template <typename T>
__global__ void k_update_history(T* nhist, const T* hist, const T* in, long long nh1, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nh1)
        return;
    const long long pos = n + i; // position in [history | input]
    nhist[i]            = pos < nh1 ? hist[pos] : in[pos - nh1];
}

template <typename T>
__global__ void k_zero(T* p, long long count)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        p[i] = T{};
}

template <typename T>
struct gpu_conv_filter<T>::impl
{
    size_t block = 1024, ntaps = 0, nh1 = 0, S = 0, L = 0;
    std::unique_ptr<typename gpu_dft_real<T>::impl> r;
    dev_mem H, hist[2], seg, spec, yseg, taps_dev;
    int cur = 0;

    impl(const T* taps, size_t nt, size_t bs)
    {
        block = next_pow2(bs == 0 ? 1024 : bs);
        set(taps, nt);
    }
    void set(const T* taps, size_t nt)
    {
        if (nt == 0)
            throw std::logic_error("convolution filter needs at least one tap");
        ntaps = nt;
        nh1   = nt - 1;
        S     = std::max<size_t>(next_pow2(std::max<size_t>(2 * block, 2 * nh1 + 2)), 2);
        L     = S - nh1;
        r.reset(new typename gpu_dft_real<T>::impl(S, 1));
        hipStream_t s = library_stream();
        taps_dev.reset(nt * sizeof(T));
        hip_check(hipMemcpy(taps_dev.ptr, taps, nt * sizeof(T), hipMemcpyHostToDevice), "filter taps upload");
        H.reset((S / 2 + 1) * sizeof(cx<T>));
        r->forward_op(load_real_pad<T>(taps_dev.ptr, 0, (long long)nt, false), H.as<cx<T>>(), 0, 1, s);
        hist[0].reset(std::max<size_t>(1, nh1) * sizeof(T));
        hist[1].reset(std::max<size_t>(1, nh1) * sizeof(T));
        hip_check(hipStreamSynchronize(s), "filter spectrum");
        reset();
    }
    void reset()
    {
        hipStream_t s = library_stream();
        for (auto& h : hist)
        {
            const long long cnt = (long long)(h.bytes / sizeof(T));
            hipLaunchKernelGGL(k_zero<T>, dim3(nblocks(cnt, 256)), dim3(256), 0, s, h.template as<T>(), cnt);
            hip_check(hipGetLastError(), "filter reset");
        }
        hip_check(hipStreamSynchronize(s), "filter reset");
        cur = 0;
    }
    void process(T* out, const T* in, size_t n, hipStream_t s)
    {
        if (n == 0)
            return;
        const long long nseg = (long long)((n + L - 1) / L);
        const long long cs   = (long long)(S / 2 + 1);
        seg.ensure((size_t)nseg * S * sizeof(T));
        spec.ensure((size_t)nseg * cs * sizeof(cx<T>));
        yseg.ensure((size_t)nseg * S * sizeof(T));
        const T* h0 = hist[cur].template as<T>();
        hipLaunchKernelGGL(k_gather_segments<T>, dim3(nblocks(nseg * (long long)S, 256)), dim3(256), 0, s,
                           seg.template as<T>(), h0, in, (long long)nh1, (long long)n, (long long)S, (long long)L,
                           nseg);
        hip_check(hipGetLastError(), "filter gather");
        r->forward_op(load_plain<T>(seg.ptr, (long long)S / 2), spec.template as<cx<T>>(), cs, nseg, s);
        r->inverse_op(spec_product<T>(spec.ptr, cs, H.template as<cx<T>>(), 0, T(1) / T(S), false),
                      store_plain<T>(yseg.ptr, (long long)S / 2), nseg, s);
        hipLaunchKernelGGL(k_scatter_valid<T>, dim3(nblocks(nseg * (long long)L, 256)), dim3(256), 0, s, out,
                           yseg.template as<T>(), (long long)nh1, (long long)n, (long long)S, (long long)L, nseg);
        hip_check(hipGetLastError(), "filter scatter");
        if (nh1 > 0)
        {
            hipLaunchKernelGGL(k_update_history<T>, dim3(nblocks((long long)nh1, 256)), dim3(256), 0, s,
                               hist[1 - cur].template as<T>(), h0, in, (long long)nh1, (long long)n);
            hip_check(hipGetLastError(), "filter history");
            cur = 1 - cur;
        }
    }
};

// complex signals and taps: the same overlap-save segments through complex S-point transforms, the
// filter spectrum FFT(h)/S premultiplied into the load of the inverse transform
template <typename T>
struct gpu_conv_filter_complex<T>::impl
{
    size_t block = 1024, ntaps = 0, nh1 = 0, S = 0, L = 0;
    std::unique_ptr<typename gpu_dft<T>::impl> f;
    dev_mem H, hist[2], seg, yseg, taps_dev;
    int cur = 0;

    impl(const void* taps, size_t nt, size_t bs)
    {
        block = next_pow2(bs == 0 ? 1024 : bs);
        set(taps, nt);
    }
    void set(const void* taps, size_t nt)
    {
        if (nt == 0)
            throw std::logic_error("convolution filter needs at least one tap");
        ntaps = nt;
        nh1   = nt - 1;
        S     = std::max<size_t>(next_pow2(std::max<size_t>(2 * block, 2 * nh1 + 2)), 2);
        L     = S - nh1;
        f.reset(new typename gpu_dft<T>::impl(S));
        hipStream_t s = library_stream();
        taps_dev.reset(nt * sizeof(cx<T>));
        hip_check(hipMemcpy(taps_dev.ptr, taps, nt * sizeof(cx<T>), hipMemcpyHostToDevice), "filter taps upload");
        H.reset(S * sizeof(cx<T>));
        xload<T> lh;
        lh.mode = ld_cpad;
        lh.p    = taps_dev.ptr;
        lh.n    = (long long)nt;
        xstore<T> sh = store_plain<T>(H.ptr, 0);
        f->run(lh, sh, 1, false, s);
        gpu_scale<T>(H.ptr, 2 * S, T(1) / T(S), s);
        hist[0].reset(std::max<size_t>(1, nh1) * sizeof(cx<T>));
        hist[1].reset(std::max<size_t>(1, nh1) * sizeof(cx<T>));
        hip_check(hipStreamSynchronize(s), "filter spectrum");
        reset();
    }
    void reset()
    {
        hipStream_t s = library_stream();
        for (auto& h : hist)
        {
            const long long cnt = (long long)(h.bytes / sizeof(cx<T>));
            hipLaunchKernelGGL(k_zero<cx<T>>, dim3(nblocks(cnt, 256)), dim3(256), 0, s, h.template as<cx<T>>(), cnt);
            hip_check(hipGetLastError(), "filter reset");
        }
        hip_check(hipStreamSynchronize(s), "filter reset");
        cur = 0;
    }
    void process(cx<T>* out, const cx<T>* in, size_t n, hipStream_t s)
    {
        if (n == 0)
            return;
        const long long nseg = (long long)((n + L - 1) / L);
        seg.ensure((size_t)nseg * S * sizeof(cx<T>));
        yseg.ensure((size_t)nseg * S * sizeof(cx<T>));
        const cx<T>* h0 = hist[cur].template as<cx<T>>();
        hipLaunchKernelGGL(k_gather_segments<cx<T>>, dim3(nblocks(nseg * (long long)S, 256)), dim3(256), 0, s,
                           seg.template as<cx<T>>(), h0, in, (long long)nh1, (long long)n, (long long)S, (long long)L,
                           nseg);
        hip_check(hipGetLastError(), "filter gather");
        f->run(load_plain<T>(seg.ptr, (long long)S), store_plain<T>(seg.ptr, (long long)S), nseg, false, s);
        xload<T> lm = load_plain<T>(seg.ptr, (long long)S);
        lm.mode     = ld_mul;
        lm.q        = H.template as<cx<T>>();
        f->run(lm, store_plain<T>(yseg.ptr, (long long)S), nseg, true, s);
        hipLaunchKernelGGL(k_scatter_valid<cx<T>>, dim3(nblocks(nseg * (long long)L, 256)), dim3(256), 0, s, out,
                           yseg.template as<cx<T>>(), (long long)nh1, (long long)n, (long long)S, (long long)L, nseg);
        hip_check(hipGetLastError(), "filter scatter");
        if (nh1 > 0)
        {
            hipLaunchKernelGGL(k_update_history<cx<T>>, dim3(nblocks((long long)nh1, 256)), dim3(256), 0, s,
                               hist[1 - cur].template as<cx<T>>(), h0, in, (long long)nh1, (long long)n);
            hip_check(hipGetLastError(), "filter history");
            cur = 1 - cur;
        }
    }
};

template <typename T>
gpu_conv_filter_complex<T>::gpu_conv_filter_complex(const void* taps_host, size_t ntaps, size_t block_size)
    : p(new impl(taps_host, ntaps, block_size))
{
}
template <typename T>
gpu_conv_filter_complex<T>::~gpu_conv_filter_complex() = default;
template <typename T>
size_t gpu_conv_filter_complex<T>::block_size() const
{
    return p->block;
}
template <typename T>
void gpu_conv_filter_complex<T>::set_data(const void* taps_host, size_t ntaps)
{
    p->set(taps_host, ntaps);
}
template <typename T>
void gpu_conv_filter_complex<T>::reset()
{
    p->reset();
}
template <typename T>
void gpu_conv_filter_complex<T>::process(void* out, const void* in, size_t n, hipStream_t s)
{
    p->process(static_cast<cx<T>*>(out), static_cast<const cx<T>*>(in), n, s);
}

template <typename T>
gpu_conv_filter<T>::gpu_conv_filter(const T* taps_host, size_t ntaps, size_t block_size)
    : p(new impl(taps_host, ntaps, block_size))
{
}
template <typename T>
gpu_conv_filter<T>::~gpu_conv_filter() = default;
template <typename T>
size_t gpu_conv_filter<T>::block_size() const
{
    return p->block;
}
template <typename T>
void gpu_conv_filter<T>::set_data(const T* taps_host, size_t ntaps)
{
    p->set(taps_host, ntaps);
}
template <typename T>
void gpu_conv_filter<T>::reset()
{
    p->reset();
}
template <typename T>
void gpu_conv_filter<T>::process(T* out, const T* in, size_t n, hipStream_t s)
{
    p->process(out, in, n, s);
}

// ------------------------------------------------------------------------------------------------
// misc

template <typename T>
__global__ void k_scale(T* p, long long count, T f)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        p[i] *= f;
}

template <typename T>
void gpu_scale(void* data, size_t count, T factor, hipStream_t s)
{
    if (!count)
        return;
    hipLaunchKernelGGL(k_scale<T>, dim3(nblocks((long long)count, 256)), dim3(256), 0, s, static_cast<T*>(data),
                       (long long)count, factor);
    hip_check(hipGetLastError(), "scale launch");
}

template class gpu_dft<float>;
template class gpu_dft<double>;
template class gpu_dft_real<float>;
template class gpu_dft_real<double>;
template class gpu_dct<float>;
template class gpu_dct<double>;
template class gpu_conv<float>;
template class gpu_conv<double>;
template class gpu_conv_filter<float>;
template class gpu_conv_filter<double>;
template class gpu_conv_filter_complex<float>;
template class gpu_conv_filter_complex<double>;
template void gpu_convolve_complex<float>(void*, const void*, size_t, const void*, size_t, bool, bool, gpu_dft<float>&,
                                          void*, void*, hipStream_t);
template void gpu_convolve_complex<double>(void*, const void*, size_t, const void*, size_t, bool, bool,
                                           gpu_dft<double>&, void*, void*, hipStream_t);
template void gpu_scale<float>(void*, size_t, float, hipStream_t);
template void gpu_scale<double>(void*, size_t, double, hipStream_t);

} // namespace kfr_amd
