// Core GPU transform machinery shared by all plans: FP64 table generation, the LDS Stockham kernels
// with fused element operations, the four-step two-pass transform for lengths beyond one workgroup,
// and the schedule (radix choice and launch geometry) for every length.
//
// Element operations connect a transform to its data: the load phase calls ld.load(b, e) for element
// e of batch item b, and the store phase calls st.store(b, e, v). Bluestein chirps, DCT mirroring,
// real-spectrum repacking and pointwise spectrum products are element operations, so none of them
// costs an extra trip through HBM. They are selected at run time (a kernel-uniform switch), which
// keeps the number of compiled kernels small.
#pragma once

#include "fft_device.hpp"
#include "fft_engine.hpp"
#include "fft_mixed.hpp"
#include "fft_pow2.hpp"
#include "fft_timing.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace kfr_amd
{

inline unsigned nblocks(long long count, int bs) { return (unsigned)((count + bs - 1) / bs); }

// kernels that use more than 64 KiB of dynamic LDS (up to the CU's 160 KiB) must say so first
template <typename K>
inline void allow_lds(K kernel, size_t bytes)
{
    if (bytes > 65536)
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes),
                  "dynamic LDS attribute");
}

// ------------------------------------------------------------------------------------------------
// element operations

enum ld_mode : int
{
    ld_plain    = 0, // p[b*pdist + e]
    ld_real_pad = 1, // real pairs p[2e], p[2e+1] of a zero-padded (optionally reversed) real signal
    ld_dct2     = 2, // (x[mirror(e)], 0)
    ld_dct3     = 3, // x[e] * W_{4n}^e (x[0] halved)
    ld_c2r      = 4, // from_fmt: packed half-length sequence from a CCs/Perm half spectrum
    ld_cpad     = 5, // zero-padded (optionally reversed) complex signal
    ld_mul      = 6  // p[b*pdist + e] * q[e] (flag: conj(q[e]))
};

enum st_mode : int
{
    st_plain      = 0, // p[b*pdist + e] = v
    st_real_trunc = 1, // real pairs, truncated to n reals, scaled
    st_dct2       = 2, // y[e] = Re(v * W_{4n}^e)
    st_dct3       = 3, // y[mirror(e)] = Re(v)
    st_ctrunc     = 4  // complex, truncated to n, scaled; flag = keep imaginary part
};

__device__ __forceinline__ long long dct_mirror(long long e, long long n)
{
    return e < (n + 1) / 2 ? 2 * e : 2 * (n - 1 - e) + 1;
}

template <typename T>
struct xload
{
    int mode      = ld_plain;
    int flag      = 0; // real_pad/cpad: reversed; c2r: pack format (1 CCs, 0 Perm); mul: conjugate q
    int src       = 0; // c2r spectrum source: 0 p, 1 p*q*scale, 2 p*conj(q)*scale
    int mconj     = 0; // conjugate the post-multiplier
    const void* p = nullptr;
    long long pdist  = 0;
    const cx<T>* q   = nullptr;
    long long qdist  = 0;
    const cx<T>* r   = nullptr; // c2r twiddles
    long long n      = 0;
    T scale          = T(1);
    long long b0     = 0;
    const cx<T>* mtab = nullptr; // post-multiplier table (Bluestein chirp)
    long long lim    = LLONG_MAX; // elements at or beyond lim are zero

    __device__ cx<T> spec(long long b, long long k) const
    {
        const cx<T> a = static_cast<const cx<T>*>(p)[b * pdist + k];
        if (src == 0)
            return a;
        const cx<T> c = q[b * qdist + k];
        return cscale(src == 1 ? cmul(a, c) : cmulc(a, c), scale);
    }

    __device__ cx<T> base(long long b, long long e) const
    {
        switch (mode)
        {
        case ld_plain:
        default:
            return static_cast<const cx<T>*>(p)[b * pdist + e];
        case ld_real_pad:
        {
            const T* x         = static_cast<const T*>(p) + b * pdist;
            const long long i0 = 2 * e, i1 = 2 * e + 1;
            T a = T(0), c = T(0);
            if (i0 < n)
                a = flag ? x[n - 1 - i0] : x[i0];
            if (i1 < n)
                c = flag ? x[n - 1 - i1] : x[i1];
            return { a, c };
        }
        case ld_dct2:
            return { static_cast<const T*>(p)[b * n + dct_mirror(e, n)], T(0) };
        case ld_dct3:
        {
            const T v = static_cast<const T*>(p)[b * n + e] * (e == 0 ? T(0.5) : T(1));
            return { v * q[e].re, v * q[e].im };
        }
        case ld_c2r:
        {
            // from_fmt (src/dft/fft-impl.hpp:1853-1896); n = half length
            const long long cs = n;
            if (e == 0)
            {
                const cx<T> x0 = spec(b, 0);
                if (flag == 1)
                {
                    const cx<T> xn = spec(b, cs);
                    return { x0.re + xn.re, x0.re - xn.re };
                }
                return { x0.re + x0.im, x0.re - x0.im };
            }
            if (2 * e == cs)
            {
                const cx<T> xm = spec(b, e);
                return { T(2) * xm.re, T(-2) * xm.im };
            }
            const bool lo     = 2 * e < cs;
            const long long k = lo ? e : cs - e;
            const cx<T> xk    = spec(b, k);
            const cx<T> xnk   = spec(b, cs - k);
            const cx<T> f1    = cadd(xk, cconj(xnk));
            const cx<T> f2    = csub(xk, cconj(xnk));
            const cx<T> t     = cmulc(f2, r[k]);
            return lo ? cadd(f1, t) : cconj(csub(f1, t));
        }
        case ld_cpad:
        {
            if (e >= n)
                return { T(0), T(0) };
            return static_cast<const cx<T>*>(p)[b * pdist + (flag ? n - 1 - e : e)];
        }
        case ld_mul:
        {
            const cx<T> a = static_cast<const cx<T>*>(p)[b * pdist + e];
            return flag ? cmulc(a, q[e]) : cmul(a, q[e]);
        }
        }
    }

    __device__ __forceinline__ cx<T> load(long long b, long long e) const
    {
        if (e >= lim)
            return { T(0), T(0) };
        cx<T> v = base(b + b0, e);
        if (mtab)
            v = mconj ? cmulc(v, mtab[e]) : cmul(v, mtab[e]);
        return v;
    }
};

template <typename T>
struct xstore
{
    int mode       = st_plain;
    int flag       = 0;
    int mconj      = 0;
    void* p        = nullptr;
    long long pdist = 0;
    const cx<T>* q = nullptr; // dct table
    long long n    = 0;
    T scale        = T(1);
    long long b0   = 0;
    const cx<T>* mtab = nullptr; // pre-multiplier (Bluestein chirp)
    long long lim  = LLONG_MAX;  // elements at or beyond lim are dropped

    __device__ void base(long long b, long long e, cx<T> v) const
    {
        switch (mode)
        {
        case st_plain:
        default:
            static_cast<cx<T>*>(p)[b * pdist + e] = v;
            break;
        case st_real_trunc:
        {
            T* y               = static_cast<T*>(p) + b * pdist;
            const long long i0 = 2 * e;
            if (i0 < n)
                y[i0] = v.re * scale;
            if (i0 + 1 < n)
                y[i0 + 1] = v.im * scale;
            break;
        }
        case st_dct2:
        {
            const cx<T> w = q[e];
            static_cast<T*>(p)[b * n + e] = v.re * w.re - v.im * w.im;
            break;
        }
        case st_dct3:
            static_cast<T*>(p)[b * n + dct_mirror(e, n)] = v.re;
            break;
        case st_ctrunc:
            if (e < n)
                static_cast<cx<T>*>(p)[b * pdist + e] = { v.re * scale, flag ? v.im * scale : T(0) };
            break;
        }
    }

    __device__ __forceinline__ void store(long long b, long long e, cx<T> v) const
    {
        if (e >= lim)
            return;
        if (mtab)
            v = mconj ? cmulc(v, mtab[e]) : cmul(v, mtab[e]);
        base(b + b0, e, v);
    }
};

template <typename T>
inline xload<T> load_plain(const void* p, long long dist)
{
    xload<T> l;
    l.p     = p;
    l.pdist = dist;
    return l;
}
template <typename T>
inline xstore<T> store_plain(void* p, long long dist)
{
    xstore<T> s;
    s.p     = p;
    s.pdist = dist;
    return s;
}

// ------------------------------------------------------------------------------------------------
// FP64 table generation: out[i] = exp(-2*pi*i*a_i/M), a_i reduced exactly in integers, evaluated
// with FP64 sincospi and rounded once to T.

enum twiddle_mode : int
{
    tw_linear = 0, // a = (i*mult + add) mod M
    tw_radix  = 1, // radix-pass table: i = k*(R-1) + (r-1), a = r*k; M = Ns*R
    tw_square = 2, // a = i^2 mod M (Bluestein chirp, M = 2N)
    tw_radix_full = 3 // radix-pass table with r = 0 included: i = k*R + r, a = r*k; M = Ns*R
};

template <typename T>
__global__ void k_twiddles(cx<T>* out, long long count, long long M, long long mult, long long add, int mode,
                           int R)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
        return;
    unsigned long long a;
    if (mode == tw_radix_full)
    {
        const long long k = i / R;
        const long long r = i - k * R;
        a                 = (unsigned long long)((r * k) % M);
    }
    else if (mode == tw_radix)
    {
        const long long k = i / (R - 1);
        const long long r = i - k * (R - 1) + 1;
        a                 = (unsigned long long)((r * k) % M);
    }
    else if (mode == tw_square)
    {
        const unsigned long long ii = (unsigned long long)i % (unsigned long long)M;
        a = (unsigned long long)(((unsigned __int128)ii * ii) % (unsigned long long)M);
    }
    else
    {
        a = (unsigned long long)(((__int128)i * mult + add) % M);
    }
    long long as = (long long)a; // symmetric range keeps the FP64 argument small
    if (2 * as > M)
        as -= M;
    const double x = 2.0 * (double)as / (double)M;
    double s, c;
    sincospi(x, &s, &c);
    out[i] = { (T)c, (T)(-s) };
}

template <typename T>
void gen_twiddles(cx<T>* out, long long count, long long M, long long mult, long long add, int mode, int R,
                  hipStream_t s)
{
    if (count <= 0)
        return;
    timing::span ts([] { return timing::kernel_name("k_twiddles", timing::type_name<T>()); },
                    (long long)sizeof(cx<T>) * count, s);
    hipLaunchKernelGGL(k_twiddles<T>, dim3(nblocks(count, 256)), dim3(256), 0, s, out, count, M, mult, add, mode,
                       R);
    hip_check(hipGetLastError(), "twiddle generation launch");
}

// ------------------------------------------------------------------------------------------------
// Stockham schedule (host): GPU radix choice, twiddle layout and launch geometry for one length.

struct stockham
{
    int n = 0;
    std::vector<int> radix, kind, ns;
    std::vector<long long> tw_off;
    long long tw_count = 0;
    int tpt            = 0;
};

inline bool lds_radix(int r) { return r == 2 || r == 3 || r == 4 || r == 5 || r == 7 || r == 8 || r == 16; }

// Pass geometry of an explicit radix list (product n): strides, twiddle offsets, threads per transform
// (wide: the ceil(pm/R) butterflies per thread of radix_pass<..., WIDE>)
inline bool stockham_from_radices(long long nn, const std::vector<int>& radices, stockham& S, int pm, bool wide = false)
{
    S = stockham{};
    if (nn < 1 || nn > (1 << 28) || (int)radices.size() > max_passes)
        return false;
    const int n = (int)nn;
    S.n         = n;
    S.radix     = radices;
    long long ns = 1, off = 0;
    int tpt      = 1;
    for (int r : S.radix)
    {
        const bool g = !lds_radix(r);
        S.kind.push_back(g ? pass_generic : pass_radix);
        S.ns.push_back((int)ns);
        S.tw_off.push_back(off);
        if (g)
        {
            off += ns * r;
            tpt = std::max(tpt, (n + pm - 1) / pm);
        }
        else
        {
            if (ns > 1)
                off += ns * (r - 1);
            const int M  = wide ? (pm + r - 1) / r : std::max(1, pm / r); // bflies_of
            const int nb = n / r;
            tpt          = std::max(tpt, (nb + M - 1) / M);
        }
        ns *= r;
    }
    if (ns != nn)
        return false;
    S.tw_count = off;
    S.tpt      = tpt;
    return true;
}

// prime factors of m other than 2 (7, 5, 3 first, then 11..127 ascending); false if another remains
inline bool odd_factors(long long& m, std::vector<int>& odd)
{
    for (int p : { 7, 5, 3 })
        while (m % p == 0)
        {
            odd.push_back(p);
            m /= p;
        }
    for (int p = 11; m > 1 && p <= 127; p += 2)
        while (m % p == 0)
        {
            odd.push_back(p);
            m /= p;
        }
    return m == 1;
}

// Factor n into GPU passes (wide: the ceil(pm/R) butterflies per thread of radix_pass<..., WIDE>).
// Power-of-2 part: the largest butterfly that fits a thread's pm values (16 for f32, 8 for f64) first,
// remainder 8/4/2; then 7, 5, 3; any other prime factor up to 127 becomes a direct DFT pass, placed last.
inline bool make_stockham(long long nn, stockham& S, int pm, bool wide = false)
{
    S = stockham{};
    if (nn < 1 || nn > (1 << 28))
        return false;
    long long m = nn;
    int a       = 0;
    while (m % 2 == 0)
    {
        m /= 2;
        ++a;
    }
    std::vector<int> radix;
    const int rbits = pm >= 16 ? 4 : 3;
    while (a >= rbits)
    {
        radix.push_back(1 << rbits);
        a -= rbits;
    }
    if (a > 0)
        radix.push_back(1 << a);
    if (!odd_factors(m, radix))
        return false;
    return stockham_from_radices(nn, radix, S, pm, wide);
}

// A finer schedule of the same transform for progressive execution, about `want` passes: every odd
// prime factor its own pass, the power-of-two part split into near-equal passes (at most 2^4 points
// for f32, 2^3 for f64, at least 2) so that the pass count approaches `want`.
inline bool make_stockham_progressive(long long nn, int want, stockham& S, int pm)
{
    long long m = nn;
    int a       = 0;
    while (m % 2 == 0)
    {
        m /= 2;
        ++a;
    }
    std::vector<int> odd;
    if (!odd_factors(m, odd))
        return false;
    const int rbits = pm >= 16 ? 4 : 3;
    int np2         = 0;
    if (a > 0)
        np2 = std::min(a, std::max((a + rbits - 1) / rbits, want - (int)odd.size()));
    std::vector<int> radix;
    for (int q = 0; q < np2; ++q)
    {
        const int bits = a / np2 + (q < a % np2 ? 1 : 0);
        radix.push_back(1 << bits);
    }
    radix.insert(radix.end(), odd.begin(), odd.end());
    return stockham_from_radices(nn, radix, S, pm);
}

template <typename T>
void gen_stockham_twiddles(const stockham& S, cx<T>* tw, hipStream_t s)
{
    for (size_t p = 0; p < S.radix.size(); ++p)
    {
        const int r  = S.radix[p];
        const int ns = S.ns[p];
        if (S.kind[p] == pass_generic)
            gen_twiddles(tw + S.tw_off[p], (long long)ns * r, (long long)ns * r, 1, 0, tw_linear, r, s);
        else if (ns > 1)
            gen_twiddles(tw + S.tw_off[p], (long long)ns * (r - 1), (long long)ns * r, 0, 0, tw_radix, r, s);
    }
}

inline lds_geom make_geom(const stockham& S, int tpb, int mapping, int ls, int ld)
{
    lds_geom G{};
    G.n       = S.n;
    G.np      = (int)S.radix.size();
    G.tpt     = S.tpt;
    G.tpb     = tpb;
    G.mapping = mapping;
    G.ls      = ls;
    G.ld      = ld;
    static const int padshift = []
    {
        const char* e = std::getenv("KFR_AMD_LDS_PADSHIFT");
        // the tiles are sized for one padding slot per 16 values (lds_padded_elems): sparser only
        return e ? std::max(4, std::min(30, atoi(e))) : 4;
    }();
    G.padshift = padshift;
    for (int p = 0; p < G.np; ++p)
    {
        G.pass[p].radix  = S.radix[p];
        G.pass[p].ns     = S.ns[p];
        G.pass[p].tw     = (int)S.tw_off[p];
        G.pass[p].kind   = S.kind[p];
        G.pass[p].ns_div = make_fastdiv((uint32_t)S.ns[p]);
    }
    return G;
}

// ------------------------------------------------------------------------------------------------
// Kernels

enum kernel_role : int
{
    role_direct = 0, // transform g is batch item g; contiguous block load and store
    role_col    = 1, // four-step pass A: column transforms, twiddled, into the intermediate
    role_row    = 2, // four-step pass B: row transforms of the intermediate, transposed store
    role_pass   = 3, // one HBM pass of a multi-pass Stockham transform: W consecutive butterflies of
                     // radix 2^L, twiddled inputs, outputs at the autosort positions
    role_conv   = 4, // a transform inside another kernel (one per workgroup, whole-length LDS)
    role_block  = 5  // a transform inside another kernel whose LDS only the passes use
};

template <typename T>
struct fourstep_args
{
    cx<T>* scratch = nullptr;
    const cx<T>* ta = nullptr; // W_n^(h*2^lbits)
    const cx<T>* tb = nullptr; // W_n^l, l < 2^lbits
    fastdiv n1div, n2div, ndiv;
    int n1 = 1, n2 = 1, lbits = 0;
    long long n = 1;
};

// Load and store phases. Each thread stages its (at most pmax) elements in registers, so that all of
// its global loads (or stores) are in flight together; the host guarantees tpb*n <= pmax*blockDim.

template <typename T>
__device__ __forceinline__ bool plain_contiguous(const xload<T>& ld, long long n)
{
    return ld.mode == ld_plain && ld.pdist == n && ld.mtab == nullptr && ld.lim == LLONG_MAX;
}
template <typename T>
__device__ __forceinline__ bool plain_contiguous(const xstore<T>& st, long long n)
{
    return st.mode == st_plain && st.pdist == n && st.mtab == nullptr && st.lim == LLONG_MAX;
}

// the block's transforms, element index fastest (e = tb*n + i); LDS layout lds[e] (ld = n, ls = 1)
template <typename T>
__device__ __forceinline__ void block_load(const lds_geom& G, cx<T>* lds, long long first, long long batch,
                                           const xload<T>& ld, const fastdiv& ndiv)
{
    const int total     = G.tpb * G.n;
    const int bs        = blockDim.x;
    const int tid       = threadIdx.x;
    const long long rem = (batch - first) * (long long)G.n;
    const int lim       = rem < total ? (int)rem : total;
    cx<T> v[pmax_of<T>];
    if (plain_contiguous(ld, G.n))
    {
        const cx<T>* src = static_cast<const cx<T>*>(ld.p) + (first + ld.b0) * (long long)G.n;
#pragma unroll
        for (int k = 0; k < pmax_of<T>; ++k)
        {
            const int e = tid + k * bs;
            v[k]        = e < lim ? src[e] : cx<T>{ T(0), T(0) };
        }
    }
    else
    {
#pragma unroll
        for (int k = 0; k < pmax_of<T>; ++k)
        {
            const int e = tid + k * bs;
            v[k]        = { T(0), T(0) };
            if (e < lim)
            {
                const int tb = (int)fdiv((uint32_t)e, ndiv);
                v[k]         = ld.load(first + tb, e - tb * G.n);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < pmax_of<T>; ++k)
    {
        const int e = tid + k * bs;
        if (e < total)
            lds[lds_pad(e, G.padshift)] = v[k];
    }
}
template <typename T>
__device__ __forceinline__ void block_store(const lds_geom& G, const cx<T>* lds, long long first, long long batch,
                                            const xstore<T>& st, const fastdiv& ndiv)
{
    const int total     = G.tpb * G.n;
    const int bs        = blockDim.x;
    const int tid       = threadIdx.x;
    const long long rem = (batch - first) * (long long)G.n;
    const int lim       = rem < total ? (int)rem : total;
    cx<T> v[pmax_of<T>];
#pragma unroll
    for (int k = 0; k < pmax_of<T>; ++k)
    {
        const int e = tid + k * bs;
        if (e < total)
            v[k] = lds[lds_pad(e, G.padshift)];
    }
    if (plain_contiguous(st, G.n))
    {
        cx<T>* dst = static_cast<cx<T>*>(st.p) + (first + st.b0) * (long long)G.n;
#pragma unroll
        for (int k = 0; k < pmax_of<T>; ++k)
        {
            const int e = tid + k * bs;
            if (e < lim)
                dst[e] = v[k];
        }
    }
    else
    {
#pragma unroll
        for (int k = 0; k < pmax_of<T>; ++k)
        {
            const int e = tid + k * bs;
            if (e < lim)
            {
                const int tb = (int)fdiv((uint32_t)e, ndiv);
                st.store(first + tb, e - tb * G.n, v[k]);
            }
        }
    }
}

template <typename T, bool INV, int ROLE, bool MIXED>
__global__ void __launch_bounds__(max_block_of(MIXED)) k_fft(lds_geom G, const cx<T>* __restrict__ tw, long long batch,
                                              xload<T> ld, xstore<T> st, fourstep_args<T> fs)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds            = reinterpret_cast<cx<T>*>(smem_raw);
    const long long first = (long long)blockIdx.x * G.tpb;
    const fft_ctx<T> c    = make_ctx<T>(G, lds, first, batch);
    cx<T> v[pmax_of<T>];
    if constexpr (ROLE == role_direct)
    {
        block_load(G, lds, first, batch, ld, fs.ndiv);
    }
    else if constexpr (ROLE == role_col)
    {
        // column col of batch item bl: elements col + n1*i; consecutive lanes take consecutive columns
        const uint32_t bl = fdiv((uint32_t)c.g, fs.n1div);
        const int col     = (int)((uint32_t)c.g - bl * (uint32_t)fs.n1);
        if (ld.mode == ld_plain && ld.mtab == nullptr && ld.lim == LLONG_MAX)
        {
            const cx<T>* src = static_cast<const cx<T>*>(ld.p) + ((long long)bl + ld.b0) * ld.pdist + col;
#pragma unroll
            for (int k = 0; k < pmax_of<T>; ++k)
            {
                const int i = c.t + k * G.tpt;
                v[k]        = (i < G.n && c.valid) ? src[(long long)fs.n1 * i] : cx<T>{ T(0), T(0) };
            }
        }
        else
        {
#pragma unroll
            for (int k = 0; k < pmax_of<T>; ++k)
            {
                const int i = c.t + k * G.tpt;
                v[k] = (i < G.n && c.valid) ? ld.load(bl, (long long)col + (long long)fs.n1 * i) : cx<T>{ T(0), T(0) };
            }
        }
#pragma unroll
        for (int k = 0; k < pmax_of<T>; ++k)
        {
            const int i = c.t + k * G.tpt;
            if (i < G.n)
                c.at(i) = v[k];
        }
    }
    else
    {
        // row g of the intermediate, contiguous
        const cx<T>* src = fs.scratch + c.g * (long long)G.n;
#pragma unroll
        for (int k = 0; k < pmax_of<T>; ++k)
        {
            const int i = c.t + k * G.tpt;
            v[k]        = (i < G.n && c.valid) ? src[i] : cx<T>{ T(0), T(0) };
        }
#pragma unroll
        for (int k = 0; k < pmax_of<T>; ++k)
        {
            const int i = c.t + k * G.tpt;
            if (i < G.n)
                c.at(i) = v[k];
        }
    }
    __syncthreads();
    fft_passes<T, INV, MIXED>(G, tw, c);
    if constexpr (ROLE == role_direct)
    {
        block_store(G, lds, first, batch, st, fs.ndiv);
    }
    else if constexpr (ROLE == role_col)
    {
        if (c.valid)
        {
            const uint32_t bl = fdiv((uint32_t)c.g, fs.n1div);
            const int col     = (int)((uint32_t)c.g - bl * (uint32_t)fs.n1);
            cx<T>* dst        = fs.scratch + (long long)bl * fs.n + col;
            const unsigned long long mask = (1ull << fs.lbits) - 1;
#pragma unroll
            for (int k = 0; k < pmax_of<T>; ++k)
            {
                const int i = c.t + k * G.tpt;
                if (i < G.n)
                    v[k] = c.at(i);
            }
#pragma unroll
            for (int k = 0; k < pmax_of<T>; ++k)
            {
                const int i = c.t + k * G.tpt;
                if (i < G.n)
                {
                    const unsigned long long m = (unsigned long long)col * (unsigned long long)i; // < n
                    const cx<T> w              = cmul(fs.ta[m >> fs.lbits], fs.tb[m & mask]);
                    dst[(long long)fs.n1 * i]  = ctw<INV>(v[k], w);
                }
            }
        }
    }
    else
    {
        // transposed store: transform index fastest, so consecutive lanes write consecutive columns k2.
        // blockDim = tpt*tpb, so a thread keeps one transform and steps i by tpt.
        const int tid    = threadIdx.x;
        const int tb     = tid % G.tpb;
        const int i0     = tid / G.tpb;
        const long long g = first + tb;
        if (g < batch)
        {
            const uint32_t bl = fdiv((uint32_t)g, fs.n2div);
            const int k2      = (int)((uint32_t)g - bl * (uint32_t)fs.n2);
#pragma unroll
            for (int k = 0; k < pmax_of<T>; ++k)
            {
                const int i = i0 + k * G.tpt;
                if (i < G.n)
                    v[k] = lds[lds_pad(tb * G.ld + i, G.padshift)];
            }
            if (st.mode == st_plain && st.mtab == nullptr && st.lim == LLONG_MAX)
            {
                cx<T>* dst = static_cast<cx<T>*>(st.p) + ((long long)bl + st.b0) * st.pdist + k2;
#pragma unroll
                for (int k = 0; k < pmax_of<T>; ++k)
                {
                    const int i = i0 + k * G.tpt;
                    if (i < G.n)
                        dst[(long long)fs.n2 * i] = v[k];
                }
            }
            else
            {
#pragma unroll
                for (int k = 0; k < pmax_of<T>; ++k)
                {
                    const int i = i0 + k * G.tpt;
                    if (i < G.n)
                        st.store(bl, (long long)k2 + (long long)fs.n2 * i, v[k]);
                }
            }
        }
    }
}

// Bluestein tables of an n-point transform padded to M points: chirp[m] = exp(-i pi (m^2 mod 2n)/n)
// (n values) and H = FFT_M(h)/M with h the chirp mirrored into both ends of M points (M values), both
// computed in FP64 and rounded once. Synchronises s.
template <typename T>
void bluestein_tables(size_t n, size_t M, cx<T>* chirp, cx<T>* H, hipStream_t s);

// ------------------------------------------------------------------------------------------------
// A transform computed without Bluestein: one LDS kernel, or four-step in two kernels.

inline size_t lds_budget() { return 128 * 1024; }

inline bool is_mixed(const stockham& s)
{
    for (int r : s.radix)
        if (r & (r - 1))
            return true;
    return false;
}

// fft_passes runs a pass only if its build has the butterfly: radix 16 where a thread holds 16 values,
// odd radices and direct prime passes in the mixed-radix build. A schedule with any other pass would
// leave it out silently, so launches check on the host first.
template <typename T>
inline void check_pass_radices(const stockham& s, bool mixed_build)
{
    for (int r : s.radix)
    {
        const bool ok = r == 2 || r == 4 || r == 8 || (r == 16 && pmax_of<T> >= 16) ||
                        (mixed_build && (r & (r - 1)) != 0 && r > 2);
        if (!ok)
            throw std::logic_error("no butterfly for a radix-" + std::to_string(r) + " pass in this kernel build");
    }
}

// launches the power-of-two or the mixed-radix build of a kernel role
template <typename T, bool INV, int ROLE>
void launch_fft(bool mixed, const lds_geom& G, const cx<T>* tw, long long batch, const xload<T>& ld,
                const xstore<T>& st, const fourstep_args<T>& fs, unsigned grid, unsigned block, size_t sh,
                hipStream_t s)
{
    // complex values in and out: batch items (columns or rows in the four-step roles) of G.n points
    timing::span ts([mixed] { return timing::kernel_name("k_fft", timing::type_name<T>(), INV, ROLE, mixed); },
                    2 * (long long)sizeof(cx<T>) * batch * G.n, s);
    if (mixed)
    {
        allow_lds(k_fft<T, INV, ROLE, true>, sh);
        hipLaunchKernelGGL((k_fft<T, INV, ROLE, true>), dim3(grid), dim3(block), sh, s, G, tw, batch, ld, st, fs);
    }
    else
    {
        allow_lds(k_fft<T, INV, ROLE, false>, sh);
        hipLaunchKernelGGL((k_fft<T, INV, ROLE, false>), dim3(grid), dim3(block), sh, s, G, tw, batch, ld, st, fs);
    }
    hip_check(hipGetLastError(), "fft launch");
}

template <typename T>
struct core_fft
{
    size_t n      = 0;
    bool fourstep = false;
    // single LDS kernel
    stockham S;
    int tpb = 1;
    // four-step: n = n1 * n2; pass A = n2-point columns (C per block), pass B = n1-point rows (R per block)
    size_t n1 = 0, n2 = 0;
    stockham SA, SB;
    int C = 1, R = 1;
    int lbits       = 0;
    long long chunk = 1; // batch items per four-step round (intermediate = chunk*n)
    dev_mem tw;
    long long twA = 0, twB = 0, twTA = 0, twTB = 0;
    dev_mem scratch;
    // power-of-two lengths and mixed lengths built from the menu: specialised kernels for contiguous
    // batches (plain loads and stores)
    std::unique_ptr<pow2_fft<T>> pow2;
    std::unique_ptr<mixed_fft<T>> mixed;

    static bool fits_lds(size_t len)
    {
        stockham s;
        if (!make_stockham((long long)len, s, pmax_of<T>))
            return false;
        return s.tpt <= max_block_of(is_mixed(s)) && len * sizeof(cx<T>) <= lds_budget();
    }

    static int pick_tpb(const stockham& s)
    {
        if (s.tpt >= 256)
            return 1;
        int t = std::max(1, 256 / s.tpt);
        // at most 64 KiB of LDS per block so that two blocks share a CU
        const size_t per = lds_padded_elems((size_t)s.n) * sizeof(cx<T>);
        t                = (int)std::max<size_t>(1, std::min<size_t>(t, 65536 / per));
        return t;
    }

    core_fft(size_t len, hipStream_t s)
    {
        n = len;
        if (pow2_fft<T>::supports(n))
            pow2.reset(new pow2_fft<T>(n, s));
        else if (mixed_fft<T>::supports(n))
            mixed.reset(new mixed_fft<T>(n, s));
        if (fits_lds(n))
        {
            make_stockham((long long)n, S, pmax_of<T>);
            check_pass_radices<T>(S, is_mixed(S));
            tpb = pick_tpb(S);
            tw.reset((size_t)std::max<long long>(1, S.tw_count) * sizeof(cx<T>));
            gen_stockham_twiddles(S, tw.as<cx<T>>(), s);
            return;
        }
        plan_fourstep();
        check_pass_radices<T>(SA, is_mixed(SA));
        check_pass_radices<T>(SB, is_mixed(SB));
        fourstep            = true;
        lbits               = (ilog2_ceil(n) + 1) / 2;
        const long long nTB = 1ll << lbits;
        const long long nTA = (long long)(n >> lbits) + 1;
        twA                 = 0;
        twB                 = SA.tw_count;
        twTA                = twB + SB.tw_count;
        twTB                = twTA + nTA;
        tw.reset((size_t)(twTB + nTB) * sizeof(cx<T>));
        cx<T>* t = tw.as<cx<T>>();
        gen_stockham_twiddles(SA, t + twA, s);
        gen_stockham_twiddles(SB, t + twB, s);
        gen_twiddles(t + twTA, nTA, (long long)n, nTB, 0, tw_linear, 0, s);
        gen_twiddles(t + twTB, nTB, (long long)n, 1, 0, tw_linear, 0, s);
        size_t cap = (size_t)256 << 20;
        if (const char* e = std::getenv("KFR_AMD_FOURSTEP_CHUNK_MB"))
            cap = (size_t)std::max(1, atoi(e)) << 20;
        chunk = (long long)std::max<size_t>(1, cap / (n * sizeof(cx<T>)));
    }

    static int ilog2_ceil(size_t v)
    {
        int l = 0;
        while ((size_t(1) << l) < v)
            ++l;
        return l;
    }

    // choose n = n1 * n2 maximising the narrower coalescing width (C columns in pass A, R rows in
    // pass B), then the balance of the two lengths
    void plan_fourstep()
    {
        size_t force_n2 = 0;
        if (const char* e = std::getenv("KFR_AMD_FOURSTEP_N2"))
            force_n2 = (size_t)atoll(e);
        size_t best_n2 = 0;
        int best_w = -1, best_C = 1, best_R = 1;
        double best_bal = 1e300;
        for (size_t d = 2; d * 2 <= n && d <= ((size_t)1 << 16); ++d)
        {
            if (n % d != 0 || (force_n2 && d != force_n2))
                continue;
            const size_t a = n / d;
            if (!fits_lds(d) || !fits_lds(a))
                continue;
            stockham sa, sb;
            make_stockham((long long)d, sa, pmax_of<T>);
            make_stockham((long long)a, sb, pmax_of<T>);
            // the columns (rows) of one workgroup may straddle two batch items: every transform computes
            // its own address and the last workgroup's surplus transforms are masked, so odd counts keep
            // full groups (one column per workgroup had made 3^10 points 1% of peak)
            int C = 16;
            const int bA = max_block_of(is_mixed(sa)), bB = max_block_of(is_mixed(sb));
            while (C > 1 && ((size_t)C * d * sizeof(cx<T>) > lds_budget() || C * sa.tpt > bA))
                C /= 2;
            int R = 16;
            while (R > 1 && ((size_t)R * (a + 1) * sizeof(cx<T>) > lds_budget() || R * sb.tpt > bB))
                R /= 2;
            const int w      = std::min(C, R);
            const double bal = std::fabs(std::log((double)a) - std::log((double)d));
            if (w > best_w || (w == best_w && bal < best_bal))
            {
                best_w   = w;
                best_bal = bal;
                best_n2  = d;
                best_C   = C;
                best_R   = R;
            }
        }
        if (!best_n2)
            throw std::runtime_error("transform length " + std::to_string(n) +
                                     " has no two-pass split on this device");
        n2 = best_n2;
        n1 = n / n2;
        C  = best_C;
        R  = best_R;
        if (const char* e = std::getenv("KFR_AMD_FOURSTEP_C"))
            C = std::max(1, std::min(C, atoi(e)));
        if (const char* e = std::getenv("KFR_AMD_FOURSTEP_R"))
            R = std::max(1, std::min(R, atoi(e)));
        make_stockham((long long)n2, SA, pmax_of<T>);
        make_stockham((long long)n1, SB, pmax_of<T>);
    }

    template <bool INV>
    void run(const xload<T>& ld, const xstore<T>& st, long long batch, hipStream_t s, workspace w = {})
    {
        if (batch <= 0)
            return;
        if ((pow2 || mixed) && plain_batch(ld) && plain_batch(st))
        {
            const cx<T>* x = static_cast<const cx<T>*>(ld.p) + ld.b0 * (long long)n;
            cx<T>* y       = static_cast<cx<T>*>(st.p) + st.b0 * (long long)n;
            if (pow2)
                pow2->run(INV, x, y, batch, s, w);
            else
                mixed->run(INV, x, y, batch, s, w);
            return;
        }
        const cx<T>* t = tw.as<cx<T>>();
        if (!fourstep)
        {
            lds_geom G      = make_geom(S, tpb, 0, 1, (int)n);
            const size_t sh = lds_padded_elems((size_t)tpb * n) * sizeof(cx<T>);
            fourstep_args<T> fs;
            fs.ndiv = make_fastdiv((uint32_t)n);
            launch_fft<T, INV, role_direct>(is_mixed(S), G, t, batch, ld, st, fs, nblocks(batch, tpb),
                                            (unsigned)(S.tpt * tpb), sh, s);
            return;
        }
        if (w.p)
        {
            run_fourstep<INV>(ld, st, batch, static_cast<cx<T>*>(w.p), 3, s,
                              rounds_in(w, chunk, n * sizeof(cx<T>), 1));
            return;
        }
        scratch.ensure((size_t)std::min<long long>(chunk, batch) * n * sizeof(cx<T>));
        run_fourstep<INV>(ld, st, batch, scratch.as<cx<T>>(), 3, s);
    }

    // four-step in rounds of at most `chunk` items through `mid`: phases bit 0 runs the column kernels,
    // bit 1 the row kernels (one of them alone only within one round)
    template <bool INV>
    void run_fourstep(const xload<T>& ld, const xstore<T>& st, long long batch, cx<T>* mid, int phases,
                      hipStream_t s, long long round = 0)
    {
        const long long chunk = round > 0 ? round : this->chunk;
        const cx<T>* t   = tw.as<cx<T>>();
        lds_geom GA      = make_geom(SA, C, 1, C, 1);
        lds_geom GB      = make_geom(SB, R, 0, 1, (int)n1 + 1);
        const size_t shA = lds_padded_elems((size_t)C * n2) * sizeof(cx<T>);
        const size_t shB = lds_padded_elems((size_t)R * (n1 + 1)) * sizeof(cx<T>);
        fourstep_args<T> fs;
        fs.scratch = mid;
        fs.ta      = t + twTA;
        fs.tb      = t + twTB;
        fs.n1div   = make_fastdiv((uint32_t)n1);
        fs.n2div   = make_fastdiv((uint32_t)n2);
        fs.n1      = (int)n1;
        fs.n2      = (int)n2;
        fs.lbits   = lbits;
        fs.n       = (long long)n;
        const bool mA = is_mixed(SA), mB = is_mixed(SB);
        for (long long b0 = 0; b0 < batch; b0 += chunk)
        {
            const long long nb = std::min(chunk, batch - b0);
            xload<T> la        = ld;
            la.b0 += b0;
            xstore<T> sb = st;
            sb.b0 += b0;
            const long long ga = nb * (long long)n1;
            fs.ndiv            = make_fastdiv((uint32_t)n2);
            if (phases & 1)
                launch_fft<T, INV, role_col>(mA, GA, t + twA, ga, la, st, fs, nblocks(ga, C), (unsigned)(SA.tpt * C),
                                             shA, s);
            const long long gb = nb * (long long)n2;
            fs.ndiv            = make_fastdiv((uint32_t)n1);
            if (phases & 2)
                launch_fft<T, INV, role_row>(mB, GB, t + twB, gb, ld, sb, fs, nblocks(gb, R), (unsigned)(SB.tpt * R),
                                             shB, s);
        }
    }

    // ---- progressive execution (one reference stage per call, include/kfr/dft/fft.hpp:393-439): one
    // item in place on `data`, as units of one launch each. Lengths in one workgroup run one pass of a
    // finer Stockham schedule per unit (about `want` passes); four-step lengths run their column and
    // row kernels as two units through `mid` (n values, kept apart from the plan's scratch).
    stockham Sp;
    dev_mem twp;
    int tpbp = 1;
    void progressive_prepare(int want, hipStream_t s)
    {
        if (fourstep || !Sp.radix.empty())
            return;
        if (!make_stockham_progressive((long long)n, std::max(1, want), Sp, pmax_of<T>))
            throw std::logic_error("no progressive schedule for this length");
        check_pass_radices<T>(Sp, is_mixed(Sp));
        if (Sp.tpt > max_block_of(is_mixed(Sp)))
            throw std::logic_error("progressive schedule needs too many threads per transform");
        tpbp = 1; // one item per progressive execution
        twp.reset((size_t)std::max<long long>(1, Sp.tw_count) * sizeof(cx<T>));
        gen_stockham_twiddles(Sp, twp.as<cx<T>>(), s);
    }
    int progressive_units() const { return fourstep ? 2 : (int)Sp.radix.size(); }
    template <bool INV>
    void progressive_unit(int u, cx<T>* data, cx<T>* mid, hipStream_t s)
    {
        if (fourstep)
        {
            run_fourstep<INV>(load_plain<T>(data, (long long)n), store_plain<T>(data, (long long)n), 1, mid,
                              u == 0 ? 1 : 2, s);
            return;
        }
        lds_geom G = make_geom(Sp, tpbp, 0, 1, (int)n);
        G.pass[0]  = G.pass[u];
        G.np       = 1;
        const size_t sh = lds_padded_elems((size_t)tpbp * n) * sizeof(cx<T>);
        fourstep_args<T> fs;
        fs.ndiv = make_fastdiv((uint32_t)n);
        launch_fft<T, INV, role_direct>(is_mixed(Sp), G, twp.as<cx<T>>(), 1, load_plain<T>(data, (long long)n),
                                        store_plain<T>(data, (long long)n), fs, 1, (unsigned)(Sp.tpt * tpbp), sh,
                                        s);
    }

    // contiguous batch: item b at b*n, no element operation
    bool plain_batch(const xload<T>& l) const
    {
        return l.mode == ld_plain && l.pdist == (long long)n && l.mtab == nullptr && l.lim == LLONG_MAX;
    }
    bool plain_batch(const xstore<T>& st) const
    {
        return st.mode == st_plain && st.pdist == (long long)n && st.mtab == nullptr && st.lim == LLONG_MAX;
    }

    std::vector<int> radices() const
    {
        if (!fourstep)
            return S.radix;
        std::vector<int> r = SA.radix;
        r.insert(r.end(), SB.radix.begin(), SB.radix.end());
        return r;
    }
};

} // namespace kfr_amd
