// Device-side building blocks of the gfx950 FFT kernels: complex arithmetic, compile-time DFT
// constants, radix-R butterflies, fast integer division, and the LDS-resident Stockham engine that
// every transform kernel is built from.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <utility>

namespace kfr_amd
{

template <typename T>
struct alignas(2 * sizeof(T)) cx
{
    T re, im;
};

template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b)
{
    return { a.re + b.re, a.im + b.im };
}
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b)
{
    return { a.re - b.re, a.im - b.im };
}
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b)
{
    return { a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re };
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> b)
{
    return { a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im };
}
// multiply by a forward twiddle (INV = false) or by its conjugate (INV = true)
template <bool INV, typename T>
__device__ __forceinline__ cx<T> ctw(cx<T> a, cx<T> w)
{
    if constexpr (INV)
        return cmulc(a, w);
    else
        return cmul(a, w);
}
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s)
{
    return { a.re * s, a.im * s };
}
template <typename T>
__device__ __forceinline__ cx<T> cconj(cx<T> a)
{
    return { a.re, -a.im };
}
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename T>
__device__ __forceinline__ cx<T> cmul_mi(cx<T> a)
{
    if constexpr (INV)
        return { -a.im, a.re };
    else
        return { a.im, -a.re };
}

// ---------------------------------------------------------------------------------------------
// Compile-time cos/sin of 2*pi*k/n (Taylor series after exact reduction of k mod n to the first
// half-turn), used for the constant twiddles inside butterflies.
constexpr double k_pi = 3.14159265358979323846264338327950288;

constexpr double taylor_sin(double x)
{
    double term = x, sum = x;
    for (int i = 1; i < 30; ++i)
    {
        term *= -x * x / ((2 * i) * (2 * i + 1));
        sum += term;
    }
    return sum;
}
constexpr double taylor_cos(double x)
{
    double term = 1, sum = 1;
    for (int i = 1; i < 30; ++i)
    {
        term *= -x * x / ((2 * i - 1) * (2 * i));
        sum += term;
    }
    return sum;
}
// cos(2*pi*k/n)
constexpr double ccos2pi(long k, long n)
{
    k %= n;
    if (k < 0)
        k += n;
    if (2 * k > n)
        k = n - k; // cos is even
    // angle in [0, pi]; fold to [0, pi/2]
    if (4 * k > n)
        return -taylor_cos(k_pi - 2.0 * k_pi * k / n);
    return taylor_cos(2.0 * k_pi * k / n);
}
// sin(2*pi*k/n)
constexpr double csin2pi(long k, long n)
{
    k %= n;
    if (k < 0)
        k += n;
    if (2 * k > n)
        return -csin2pi(n - k, n);
    if (4 * k > n)
        return taylor_sin(k_pi - 2.0 * k_pi * k / n);
    return taylor_sin(2.0 * k_pi * k / n);
}

// ---------------------------------------------------------------------------------------------
// Butterflies: in-place DFT of R points, X[k] = sum_n x[n] exp(-+2*pi*i*n*k/R)
// (minus sign for the forward transform, plus for the inverse). No scaling.

template <bool INV, typename T>
__device__ __forceinline__ void bfly2(cx<T>& a, cx<T>& b)
{
    cx<T> t = a;
    a       = cadd(t, b);
    b       = csub(t, b);
}

template <bool INV, typename T>
__device__ __forceinline__ void bfly4(cx<T>& a0, cx<T>& a1, cx<T>& a2, cx<T>& a3)
{
    cx<T> t0 = cadd(a0, a2), t1 = csub(a0, a2);
    cx<T> t2 = cadd(a1, a3), t3 = cmul_mi<INV>(csub(a1, a3));
    a0       = cadd(t0, t2);
    a2       = csub(t0, t2);
    a1       = cadd(t1, t3);
    a3       = csub(t1, t3);
}

// multiply by exp(-+2*pi*i*k/n) with compile-time k, n
template <bool INV, long K, long N, typename T>
__device__ __forceinline__ cx<T> twc(cx<T> a)
{
    constexpr long k = ((K % N) + N) % N;
    if constexpr (k == 0)
        return a;
    else if constexpr (4 * k == N)
        return cmul_mi<INV>(a);
    else if constexpr (2 * k == N)
        return { -a.re, -a.im };
    else if constexpr (4 * k == 3 * N)
        return cmul_mi<!INV>(a);
    else
    {
        constexpr T c = T(ccos2pi(k, N));
        constexpr T s = INV ? T(csin2pi(k, N)) : T(-csin2pi(k, N));
        return { a.re * c - a.im * s, a.re * s + a.im * c };
    }
}

template <bool INV, typename T>
__device__ __forceinline__ void bfly8(cx<T>* v)
{
    // radix-2 x radix-4: 4-point DFTs over even/odd, combined with W8 twiddles
    bfly4<INV>(v[0], v[2], v[4], v[6]);
    bfly4<INV>(v[1], v[3], v[5], v[7]);
    v[3] = twc<INV, 1, 8>(v[3]);
    v[5] = twc<INV, 2, 8>(v[5]);
    v[7] = twc<INV, 3, 8>(v[7]);
    bfly2<INV>(v[0], v[1]);
    bfly2<INV>(v[2], v[3]);
    bfly2<INV>(v[4], v[5]);
    bfly2<INV>(v[6], v[7]);
    // outputs now at: X0=v0, X4=v1, X1=v2, X5=v3, X2=v4, X6=v5, X3=v6, X7=v7
    cx<T> x[8] = { v[0], v[2], v[4], v[6], v[1], v[3], v[5], v[7] };
#pragma unroll
    for (int i = 0; i < 8; ++i)
        v[i] = x[i];
}

template <bool INV, typename T>
__device__ __forceinline__ void bfly16(cx<T>* v)
{
    // 4 x 4: column DFTs over n = n1 + 4*n2 (stride-4 groups), twiddle W16^(n1*k2), row DFTs
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1)
        bfly4<INV>(v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]);
    // element (n1, k2) sits at v[n1 + 4*k2]
    v[5]  = twc<INV, 1, 16>(v[5]);
    v[6]  = twc<INV, 2, 16>(v[6]);
    v[7]  = twc<INV, 3, 16>(v[7]);
    v[9]  = twc<INV, 2, 16>(v[9]);
    v[10] = twc<INV, 4, 16>(v[10]);
    v[11] = twc<INV, 6, 16>(v[11]);
    v[13] = twc<INV, 3, 16>(v[13]);
    v[14] = twc<INV, 6, 16>(v[14]);
    v[15] = twc<INV, 9, 16>(v[15]);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2)
        bfly4<INV>(v[4 * k2], v[4 * k2 + 1], v[4 * k2 + 2], v[4 * k2 + 3]);
    // (k2, k1) at v[4*k2 + k1] holds X[k2 + 4*k1]
    cx<T> x[16];
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2)
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1)
            x[k2 + 4 * k1] = v[4 * k2 + k1];
#pragma unroll
    for (int i = 0; i < 16; ++i)
        v[i] = x[i];
}

// An opaque copy of v (an empty asm statement that the compiler must assume rewrites the register).
// Indices derived from the copy are not merged with those derived from v, so a kernel recomputes its
// cheap per-slot offsets in each phase instead of keeping dozens of them live (and spilled) across
// the phases in between.
__device__ __forceinline__ int fresh(int v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// Makes a value exist in registers at this point: the arithmetic that forms it cannot sink below,
// so the (more numerous) operands it was formed from die here.
template <typename T>
__device__ __forceinline__ void settle(cx<T>& a)
{
    asm volatile("" : "+v"(a.re), "+v"(a.im));
}

// f(std::integral_constant<int, I>) for I = 0 .. N-1, unrolled at compile time
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
    [&]<int... I>(std::integer_sequence<int, I...>) { (f(std::integral_constant<int, I>{}), ...); }(
        std::make_integer_sequence<int, N>{});
}

// Odd prime R: conjugate-pair form. X[0] = sum; for k>0 X[k] = x0 + sum_{n=1}^{h}
// (x_n + x_{R-n}) cos(2 pi n k/R) -+ i (x_n - x_{R-n}) sin(2 pi n k/R).
template <int R, bool INV, typename T>
__device__ __forceinline__ void bfly_odd(cx<T>* v)
{
    constexpr int h = (R - 1) / 2;
    cx<T> s[h], d[h];
    cx<T> x0 = v[0], sum = v[0];
#pragma unroll
    for (int n = 1; n <= h; ++n)
    {
        s[n - 1] = cadd(v[n], v[R - n]);
        d[n - 1] = csub(v[n], v[R - n]);
        sum      = cadd(sum, s[n - 1]);
    }
    v[0] = sum;
    // the cosines and sines are template constants: evaluated by the compiler, never at run time
    static_for<h>([&](auto km1)
    {
        constexpr int k = decltype(km1)::value + 1;
        cx<T> a = x0, b = { T(0), T(0) };
        static_for<h>([&](auto nm1)
        {
            constexpr int n  = decltype(nm1)::value + 1;
            constexpr T c    = T(ccos2pi(long(n) * k, R));
            constexpr T sn   = T(csin2pi(long(n) * k, R));
            a.re += s[n - 1].re * c;
            a.im += s[n - 1].im * c;
            b.re += d[n - 1].re * sn;
            b.im += d[n - 1].im * sn;
        });
        // forward: X[k] = a - i b, X[R-k] = a + i b ; inverse: swapped signs
        cx<T> ib = { -b.im, b.re };
        if constexpr (INV)
        {
            v[k]     = cadd(a, ib);
            v[R - k] = csub(a, ib);
        }
        else
        {
            v[k]     = csub(a, ib);
            v[R - k] = cadd(a, ib);
        }
    });
}

template <int R, bool INV, typename T>
__device__ __forceinline__ void butterfly(cx<T>* v)
{
    if constexpr (R == 2)
        bfly2<INV>(v[0], v[1]);
    else if constexpr (R == 4)
        bfly4<INV>(v[0], v[1], v[2], v[3]);
    else if constexpr (R == 8)
        bfly8<INV>(v);
    else if constexpr (R == 16)
        bfly16<INV>(v);
    else
        bfly_odd<R, INV>(v);
}

// ---------------------------------------------------------------------------------------------
// Division by a runtime-invariant divisor d < 2^31 for dividends n < 2^31:
// q = (umulhi(n, m) + n) >> s with s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1.
struct fastdiv
{
    uint32_t d = 1, m = 1, s = 0;
};

__host__ inline fastdiv make_fastdiv(uint32_t d)
{
    fastdiv f;
    f.d      = d;
    uint32_t s = 0;
    while ((uint64_t(1) << s) < d)
        ++s;
    f.s = s;
    f.m = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << s) - d)) / d + 1);
    return f;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const fastdiv& f)
{
    return (uint32_t)(((uint64_t)__umulhi(n, f.m) + n) >> f.s);
}

// ---------------------------------------------------------------------------------------------
// LDS Stockham engine.
//
// A transform of length n is computed by passes p = 0..np-1 of radix R_p. With Ns the product of the
// radices before pass p, butterfly j (0 <= j < n/R) reads x[j + r*n/R] (r = 0..R-1), multiplies
// element r by W_{Ns*R}^{r*(j mod Ns)}, runs the R-point DFT and writes element r to
// (j - j mod Ns)*R + j mod Ns + r*Ns. The last pass has Ns*R = n and writes j + r*n/R, so the
// result is in natural order without a reorder pass. The transform is loaded into LDS with
// coalesced reads, every pass runs LDS -> registers -> LDS, and the result is stored from LDS, so
// every HBM element is touched once.

constexpr int max_passes = 20;
// Complex values a thread holds per pass and stages per load/store phase: 16 complex f32 (a radix-16
// butterfly) or 8 complex f64 (radix 8). Kernels come in two register classes: power-of-two radices
// only, with up to 1024 threads per block (128 VGPRs each), and mixed radices (odd butterflies and
// direct prime passes need more temporaries), with up to 512 threads (256 VGPRs each). Both fit
// without spilling to scratch.
// The f64 value is a build option so that the CPU emulation of these kernels (tests/cpu_emu) can check
// the 16-points-per-thread geometry for f64 too; the library is built with 8.
#ifndef KFR_AMD_PMAX_F64
#define KFR_AMD_PMAX_F64 8
#endif
template <typename T>
inline constexpr int pmax_of = sizeof(T) == 4 ? 16 : KFR_AMD_PMAX_F64;
constexpr int max_block_pow2  = 1024;
constexpr int max_block_mixed = 512;
__host__ __device__ constexpr int max_block_of(bool mixed) { return mixed ? max_block_mixed : max_block_pow2; }

enum pass_kind : int
{
    pass_radix   = 0, // specialised butterfly, R in {2,3,4,5,7,8,16}
    pass_generic = 1  // direct R-point DFT (R prime, 11..127) from a W_{Ns*R} table
};

struct pass_desc
{
    int radix;
    int ns;
    int tw; // offset of this pass's twiddles in the plan's table (complex elements)
    int kind;
    fastdiv ns_div;
};

struct lds_geom
{
    int n;       // transform length
    int np;      // number of passes
    int tpt;     // threads per transform
    int tpb;     // transforms per block
    int mapping; // 0: transform = tid / tpt ; 1: transform = tid % tpb
    int ls;      // LDS element stride
    int ld;      // LDS distance between transforms
    int padshift; // LDS padding: one slot per 2^padshift elements
    pass_desc pass[max_passes];
};

// LDS layout: logical element x lives at x + x/16, one padding slot per 16 elements, so that the
// power-of-two strides of the Stockham passes (stride-R writes of the first pass, one transform per
// lane for short lengths) spread over the banks instead of hitting one.
__device__ __forceinline__ int lds_pad(int x, int s) { return x + (x >> s); }
__host__ __device__ inline size_t lds_padded_elems(size_t logical) { return logical + logical / 16 + 1; }

template <typename T>
struct fft_ctx
{
    cx<T>* lds;  // the block's LDS
    int base;    // this transform's logical LDS offset (tb * ld)
    int ls;      // logical element stride
    int ps;      // padding shift
    int t;       // thread index within the transform
    int tb;      // transform index within the block
    bool valid;  // transform exists (batch tail)
    long long g; // global transform index
    // element i of this transform (padded layout, see lds_pad)
    __device__ __forceinline__ cx<T>& at(int i) const { return lds[lds_pad(base + i * ls, ps)]; }
};

// butterflies per thread in a radix-R pass: floor(pmax/R) (at least 1), or with WIDE ceil(pmax/R), so
// that R*M >= pmax and every length up to 512*pmax runs on at most 512 threads per transform
template <typename T, int R, bool WIDE>
inline constexpr int bflies_of = WIDE ? (pmax_of<T> + R - 1) / R : (pmax_of<T> / R > 0 ? pmax_of<T> / R : 1);

// one radix pass, LDS -> LDS
template <typename T, int R, bool INV, bool WIDE = false>
__device__ __forceinline__ void radix_pass(const lds_geom& G, const pass_desc& P, const cx<T>* __restrict__ tw,
                                           const fft_ctx<T>& c)
{
    constexpr int M = bflies_of<T, R, WIDE>;
    const int nb    = G.n / R;
    cx<T> v[M][R];
    int kk[M];
#pragma unroll
    for (int m = 0; m < M; ++m)
    {
        const int j = c.t + m * G.tpt;
        if (j < nb)
        {
#pragma unroll
            for (int r = 0; r < R; ++r)
                v[m][r] = c.at(j + r * nb);
        }
    }
#pragma unroll
    for (int m = 0; m < M; ++m)
    {
        const int j = c.t + m * G.tpt;
        kk[m]       = 0;
        if (j < nb)
        {
            const int k = P.ns == 1 ? 0 : (int)(j - fdiv(j, P.ns_div) * P.ns);
            kk[m]       = k;
            if (P.ns > 1)
            {
                const cx<T>* w = tw + P.tw + k * (R - 1);
#pragma unroll
                for (int r = 1; r < R; ++r)
                    v[m][r] = ctw<INV>(v[m][r], w[r - 1]);
            }
            butterfly<R, INV>(v[m]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < M; ++m)
    {
        const int j = c.t + m * G.tpt;
        if (j < nb)
        {
            const int base = (j - kk[m]) * R + kk[m];
#pragma unroll
            for (int r = 0; r < R; ++r)
                c.at(base + r * P.ns) = v[m][r];
        }
    }
    __syncthreads();
}

// direct DFT pass for a prime radix without a specialised butterfly, LDS -> LDS.
// Output o = q*Ns*R + k2*Ns + k is sum_r x[j + r*n/R] W_{Ns*R}^{r*(k + k2*Ns)} with j = q*Ns + k;
// the pass's table holds W_{Ns*R}^m for m in [0, Ns*R).
template <typename T, bool INV>
__device__ __forceinline__ void generic_pass(const lds_geom& G, const pass_desc& P, const cx<T>* __restrict__ tw,
                                             const fft_ctx<T>& c)
{
    const int R  = P.radix;
    const int nb = G.n / R;
    const int M  = P.ns * R; // table length
    const cx<T>* w = tw + P.tw;
    cx<T> out[pmax_of<T>];
#pragma unroll
    for (int m = 0; m < pmax_of<T>; ++m)
    {
        const int o = c.t + m * G.tpt;
        out[m]      = { T(0), T(0) };
        if (o < G.n)
        {
            const int q    = o / M;
            const int rem  = o - q * M;
            const int k2   = rem / P.ns;
            const int k    = rem - k2 * P.ns;
            const int j    = q * P.ns + k;
            const int step = k + k2 * P.ns;
            int idx        = 0;
            cx<T> acc      = { T(0), T(0) };
            for (int r = 0; r < R; ++r)
            {
                acc = cadd(acc, ctw<INV>(c.at(j + r * nb), w[idx]));
                idx += step;
                if (idx >= M)
                    idx -= M;
            }
            out[m] = acc;
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < pmax_of<T>; ++m)
    {
        const int o = c.t + m * G.tpt;
        if (o < G.n)
            c.at(o) = out[m];
    }
    __syncthreads();
}

// all passes of the transform held in LDS (natural order in, natural order out). MIXED = false
// compiles only the power-of-two butterflies (radix 16 for f32, 8 for f64, and the remainders). WIDE
// selects the ceil(pmax/R) butterflies per thread of radix_pass (the geometry must come from
// make_stockham with wide = true).
template <typename T, bool INV, bool MIXED, bool WIDE = false>
__device__ __forceinline__ void fft_passes(const lds_geom& G, const cx<T>* tw, const fft_ctx<T>& c)
{
    for (int p = 0; p < G.np; ++p)
    {
        const pass_desc& P = G.pass[p];
        switch (P.radix)
        {
        case 2:
            radix_pass<T, 2, INV, WIDE>(G, P, tw, c);
            break;
        case 4:
            radix_pass<T, 4, INV, WIDE>(G, P, tw, c);
            break;
        case 8:
            radix_pass<T, 8, INV, WIDE>(G, P, tw, c);
            break;
        case 16: // only where a thread holds 16 values (check_pass_radices rejects it elsewhere)
            if constexpr (pmax_of<T> >= 16)
                radix_pass<T, 16, INV, WIDE>(G, P, tw, c);
            break;
        default:
            if constexpr (MIXED)
            {
                switch (P.radix)
                {
                case 3:
                    radix_pass<T, 3, INV, WIDE>(G, P, tw, c);
                    break;
                case 5:
                    radix_pass<T, 5, INV, WIDE>(G, P, tw, c);
                    break;
                case 7:
                    radix_pass<T, 7, INV, WIDE>(G, P, tw, c);
                    break;
                default:
                    generic_pass<T, INV>(G, P, tw, c);
                    break;
                }
            }
            break;
        }
    }
}

template <typename T>
__device__ __forceinline__ fft_ctx<T> make_ctx(const lds_geom& G, cx<T>* lds, long long first_g, long long batch)
{
    fft_ctx<T> c;
    const int tid = threadIdx.x;
    if (G.mapping == 0)
    {
        c.tb = tid / G.tpt;
        c.t  = tid - c.tb * G.tpt;
    }
    else
    {
        c.t  = tid / G.tpb;
        c.tb = tid - c.t * G.tpb;
    }
    c.g     = first_g + c.tb;
    c.valid = c.g < batch;
    c.lds   = lds;
    c.base  = c.tb * G.ld;
    c.ls    = G.ls;
    c.ps    = G.padshift;
    return c;
}

} // namespace kfr_amd
