// Power-of-two C2C kernels specialised per length.
//
// A transform of N = 2^L points is split over TPT = N/E threads (E = 16 points per thread for f32,
// 8 for f64, twice that where a workgroup would otherwise need more than 512 threads). Every pass of
// the Stockham autosort schedule reads, in thread t, the strided set x[t + k*TPT] (k < E), so the
// first pass takes its operands straight from HBM (coalesced across threads), the passes exchange
// data through LDS (in two rounds of half the positions when E is doubled), and the last pass, whose
// outputs are the same strided set in natural order, writes straight to HBM. Radix-16 (f32) or
// radix-8 (f64) passes follow one smaller power-of-two pass that absorbs the remainder; with the
// radices, pass strides, table offsets and loop bounds all compile-time constants, the butterflies
// run from registers, and each pass builds its twiddles from the power-of-two entries of its table
// row.
//
// Lengths beyond one workgroup's LDS use the four-step split n = n1*n2 with x[c + n1*i]:
// kernel "col" transforms the n2-point columns (W columns per workgroup, consecutive lanes on
// consecutive columns), multiplies by W_n^(c*k2) (built from table powers) and writes the
// intermediate; kernel "row"
// transforms the n1-point rows and stores them transposed through LDS, so both kernels move
// W-element runs. Workgroups of both kernels are ordered so that neighbouring column (or row)
// groups land on the same XCD and share its L2.
//
// The longest lengths run as a multi-pass Stockham over HBM instead: passes of radix 2^5..2^8, each
// one kernel ("pass") in which a workgroup transforms 16 butterflies of stride Ns (consecutive lanes
// on consecutive butterflies), with input twiddles W_{Ns*R}^(r*(j mod Ns)) from a two-level table,
// and writes outputs (j - j mod Ns)*R + j mod Ns + r*Ns; the first pass (Ns = 1) writes its
// contiguous output runs through LDS.
//
// Half-precision storage (bfloat16 or binary16 pairs, T = float) reuses the same kernels with
// conversions at load and store; their radix-16 passes run as 16-point DFT-matrix products on MFMA.
#include "fft_pow2.hpp"
#include "fft_timing.hpp"

#include "fft_core.hpp"

#include <algorithm>
#include <stdexcept>

namespace kfr_amd
{
namespace p2
{

template <typename T>
inline constexpr int log_e = sizeof(T) == 4 ? 4 : 3; // log2 points per thread
template <typename T>
inline constexpr int max_log = sizeof(T) == 4 ? 14 : 13; // longest four-step factor
// longest full-precision transform in one workgroup: one length beyond the four-step factors, with
// two exchange rounds (half of the transform in LDS at a time, 1024 threads, 136 KiB)
template <typename T>
inline constexpr int direct_max_log = max_log<T> + 1;
template <typename T>
inline constexpr int lds_budget = sizeof(T) == 4 ? 16384 : 8192; // complex values per workgroup
constexpr int min_split_log = 7;                                 // shortest four-step factor

// pass schedule of 2^L points (also used on the host to lay out the twiddle tables)
__host__ __device__ constexpr int pass_le(int L, int le) { return L < le ? L : le; }
__host__ __device__ constexpr int pass_rem(int L, int le) { return L % pass_le(L, le); }
__host__ __device__ constexpr int pass_count(int L, int le)
{
    return L / pass_le(L, le) + (pass_rem(L, le) ? 1 : 0);
}
__host__ __device__ constexpr int pass_lr(int L, int le, int p) // log2 radix of pass p
{
    return (pass_rem(L, le) && p == 0) ? pass_rem(L, le) : pass_le(L, le);
}
__host__ __device__ constexpr int pass_lns(int L, int le, int p) // log2 of the stride Ns of pass p
{
    int s = 0;
    for (int q = 0; q < p; ++q)
        s += pass_lr(L, le, q);
    return s;
}
// pass p's table holds W_{Ns*R}^(r*k) at k*R + r (k < Ns, r < R); passes with Ns = 1 have none
__host__ __device__ constexpr int pass_twoff(int L, int le, int p)
{
    int o = 0;
    for (int q = 0; q < p; ++q)
        if (pass_lns(L, le, q) > 0)
            o += 1 << (pass_lns(L, le, q) + pass_lr(L, le, q));
    return o;
}

// ceil(log2(v)) of a positive compile-time value
__host__ __device__ constexpr int ilog2c(int v)
{
    int l = 0;
    while ((1 << l) < v)
        ++l;
    return l;
}

// EX: points per thread as a multiple of the radix of the passes (2 gives every thread two
// butterflies per pass and halves the threads per transform)
template <typename T, int L, int EX = 1>
struct shape
{
    static constexpr int N   = 1 << L;
    static constexpr int LE  = pass_le(L, log_e<T>);
    static constexpr int E   = (1 << LE) * EX;
    static constexpr int TPT = N / E;
    static constexpr int NP  = pass_count(L, log_e<T>);
    static constexpr int lr(int p) { return pass_lr(L, log_e<T>, p); }
    static constexpr int lns(int p) { return pass_lns(L, log_e<T>, p); }
    static constexpr int twoff(int p) { return pass_twoff(L, log_e<T>, p); }
    static constexpr int twcount = pass_twoff(L, log_e<T>, pass_count(L, log_e<T>));
};

template <typename T, int L, int ROLE, int PREC = 0>
struct geom
{
    using S1 = shape<T, L, 1>;
    static constexpr int width()
    {
        if (ROLE == role_conv || ROLE == role_block)
            return 1;
        if (ROLE == role_direct)
            return S1::TPT >= 256 ? 1 : 256 / S1::TPT;
        // full-precision rows that would exceed the budget exchange in two rounds (ex() below), holding
        // half of each row in LDS at a time: twice the rows per workgroup and twice the run length of
        // the transposed store (measured 2-6% faster for rows of 2^11..2^13 points); the same for the
        // columns of which the budget holds fewer than 16 (f32 2^23 2.5%, f64 2^22 and 2^23 4-7% faster)
        constexpr int grow = (ROLE == role_row || (ROLE == role_col && 16 * S1::N > lds_budget<T>)) && PREC == 0 &&
                                     S1::LE >= log_e<T>
                                 ? 2
                                 : 1;
        // half-precision passes hold 32 butterflies: 128-byte runs of 4-byte values
        int w = ROLE == role_pass && PREC != 0 ? 32 : 16;
        while (w > 1 && (w * S1::N > grow * lds_budget<T> || w * S1::TPT > grow * 1024))
            w /= 2;
        // runs longer than a 128-byte line buy nothing, two workgroups per CU do (measured for f64 at
        // 2^9 points: 512 threads and 128-byte runs beat 1024 threads and 256-byte runs)
        if (ROLE != role_pass && w * (int)sizeof(cx<T>) > 128 && w * S1::TPT > 512)
            w /= 2;
        return w;
    }
    static constexpr int W = width(); // transforms per workgroup
    // Workgroups that would need more than 512 threads (and so more than 64 KiB of LDS) give every
    // thread twice the points and exchange through LDS in two halves: same runs in HBM, half the
    // threads and half the LDS, so two workgroups share a CU.
    static constexpr int ex()
    {
        // measured: column kernels of which the budget holds 16 columns (their strided loads, twiddles
        // and stores hold the most registers) and the half-precision kernels (MFMA operands) run faster
        // with the full workgroup
        if (ROLE == role_pass || ROLE == role_conv || PREC != 0 || S1::LE < log_e<T> ||
            (ROLE == role_col && 16 * S1::N <= lds_budget<T>))
            return 1;
        if (ROLE == role_direct || ROLE == role_block) // measured: f64 from 128 threads, f32 from 512
            return S1::TPT >= (sizeof(T) == 8 ? 128 : 512) ? 2 : 1;
        return W * S1::TPT > 512 ? 2 : 1;
    }
    static constexpr int EX = ex();
    static constexpr int H  = EX; // LDS exchange rounds
    using S                 = shape<T, L, EX>;
    static constexpr int BS = S::TPT * W;
    static constexpr int NH = S::N / H; // positions per exchange round
    static constexpr int LD = ROLE == role_row ? NH + 1 : NH; // LDS distance between transforms
    static constexpr int logical = ROLE == role_row ? W * LD : W * NH;
    static constexpr size_t lds_bytes = (size_t)(logical + logical / 16 + 1) * sizeof(cx<T>);
    // waves per SIMD that the workgroups' LDS lets one CU hold, up to 4 (128 VGPRs): the one-workgroup
    // kernels (plain, real and DCT) and the row kernels are compiled for it so that their registers do
    // not cost a workgroup per CU
    static constexpr int lds_waves = std::min<int>(4, (int)(163840 / lds_bytes) * (BS / 64) / 4);
    // short transforms (under 16 threads each) load and store through LDS in whole-workgroup runs
    static constexpr bool staged = ROLE == role_direct && S::TPT < 16;
    static_assert(BS <= 1024, "workgroup too large");
    static_assert(H == 1 || !staged, "two-round exchange needs whole-thread transforms");
};

// the values one thread holds
template <typename T, int L, int ROLE, int PREC>
using regs = cx<T>[geom<T, L, ROLE, PREC>::S::E];

// LDS slot of element i of the workgroup's transform tb: columns are interleaved (consecutive lanes
// on consecutive columns), other roles keep each transform contiguous; one padding slot per 16
// values spreads the power-of-two strides of the passes over the banks.
template <int ROLE, int W, int LD>
__device__ __forceinline__ int laddr(int tb, int i)
{
    const int x = (ROLE == role_col || ROLE == role_pass) ? i * W + tb : tb * LD + i;
    return x + (x >> 4);
}

template <typename T, int R>
__device__ __forceinline__ void load_table(const cx<T>* p, cx<T> (&w)[R])
{
    if constexpr (sizeof(T) == 4 && R >= 2)
    {
        const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
        for (int i = 0; i < R / 2; ++i)
        {
            const float4 f = q[i];
            w[2 * i]       = { f.x, f.y };
            w[2 * i + 1]   = { f.z, f.w };
        }
    }
    else
    {
#pragma unroll
        for (int i = 0; i < R; ++i)
            w[i] = p[i];
    }
}

template <typename T>
struct args
{
    const void* in   = nullptr;
    void* out        = nullptr;
    const cx<T>* tw  = nullptr; // pass tables of this kernel's length
    const cx<T>* ta  = nullptr; // four-step: W_n^(h*2^lbits)
    const cx<T>* tb  = nullptr; // four-step: W_n^l, l < 2^lbits
    int* scales      = nullptr; // half-precision four-step: power-of-two scale of each column group
    long long count  = 0;       // transforms
    int lo           = 0;       // four-step: log2 of the other dimension (col: n1, row: n2)
    int lbits        = 0;
    int gshift       = 0; // half-precision row kernel: log2 of the column kernel's columns per workgroup
    int ns_log       = 0; // multi-pass: log2 of the stride Ns of this pass (twiddles W_{Ns*R} from ta, tb)
    unsigned nblk    = 0; // workgroups with work (the grid of split kernels is a multiple of 8)
    // half-precision multi-pass: the input is the previous pass's binary16 intermediate, element o of
    // item b scaled by in_scales[workgroup of the previous pass that wrote it]; that pass had radix
    // 2^in_l, stride 2^in_ns and 2^in_wlog butterflies per workgroup
    const int* in_scales = nullptr;
    int in_l = 0, in_ns = 0, in_wlog = 0;
    int out_mid = 0; // write a binary16 intermediate with one power-of-two scale per workgroup (a.scales)
};

// ---- storage precision of the data in HBM (the arithmetic is T throughout)
//   prec_full: cx<T>; prec_bf16 / prec_f16: complex values of two bfloat16 / binary16 (T = float only).
// The half-precision four-step keeps its intermediate as binary16 with one power-of-two scale per
// column-kernel workgroup, so values of any range keep 11-bit precision.
enum storage_prec : int
{
    prec_full = 0,
    prec_bf16 = 1,
    prec_f16  = 2
};

template <typename H>
struct alignas(2 * sizeof(H)) hcx
{
    H re, im;
};

template <typename T, int PREC>
__device__ __forceinline__ cx<T> ld_elem(const void* p, long long i)
{
    if constexpr (PREC == prec_full)
        return static_cast<const cx<T>*>(p)[i];
    else if constexpr (PREC == prec_bf16)
    {
        const hcx<__bf16> h = static_cast<const hcx<__bf16>*>(p)[i];
        return { (T)(float)h.re, (T)(float)h.im };
    }
    else
    {
        const hcx<_Float16> h = static_cast<const hcx<_Float16>*>(p)[i];
        return { (T)(float)h.re, (T)(float)h.im };
    }
}

// NT: a nontemporal store, for kernels whose stores fill whole 128-byte lines (the four-step, multi-pass
// and f64 one-workgroup lengths measured 1-5% faster with it; kernels that write 64-byte runs, whose
// half lines merge in L2, up to 12% slower)
template <typename T, int PREC, bool NT = false>
__device__ __forceinline__ void st_elem(void* p, long long i, cx<T> v)
{
    if constexpr (PREC == prec_full && NT)
    {
        using V = T __attribute__((ext_vector_type(2)));
        __builtin_nontemporal_store(V{ v.re, v.im }, reinterpret_cast<V*>(static_cast<cx<T>*>(p) + i));
    }
    else if constexpr (PREC == prec_full)
        static_cast<cx<T>*>(p)[i] = v;
    else if constexpr (PREC == prec_bf16)
        static_cast<hcx<__bf16>*>(p)[i] = { (__bf16)(float)v.re, (__bf16)(float)v.im };
    else
        static_cast<hcx<_Float16>*>(p)[i] = { (_Float16)(float)v.re, (_Float16)(float)v.im };
}

// precision of the four-step intermediate
template <int PREC>
inline constexpr int mid_prec = PREC == prec_full ? prec_full : prec_f16;

// ---- radix-16 butterflies on MFMA (half-precision storage): the 16-point DFT of 32 butterflies is
// the real product [Re y; Im y] = [[Re F, -Im F], [Im F, Re F]] [Re x; Im x] with the 32x32 matrix as
// the A operand of v_mfma_f32_32x32x16_bf16 (lane l, slot s: A[l%32][8(l/32)+s]; B[8(l/32)+s][l%32];
// D slot v: D[8(v/4)+4(l/32)+v%4][l%32]). Matrix and data are each split into bf16 hi + lo parts and
// three products are accumulated in FP32 (hi*hi + hi*lo + lo*hi), about 16 significant bits.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

struct mfma16
{
    bf16x8 a0h, a0l, a1h, a1l; // chunk 0 multiplies Re x, chunk 1 Im x
};

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo)
{
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

template <bool INV>
__device__ __forceinline__ mfma16 make_mfma16()
{
    // cos and sin of 2*pi*q/16
    constexpr float cq[16] = { (float)ccos2pi(0, 16),  (float)ccos2pi(1, 16),  (float)ccos2pi(2, 16),
                               (float)ccos2pi(3, 16),  (float)ccos2pi(4, 16),  (float)ccos2pi(5, 16),
                               (float)ccos2pi(6, 16),  (float)ccos2pi(7, 16),  (float)ccos2pi(8, 16),
                               (float)ccos2pi(9, 16),  (float)ccos2pi(10, 16), (float)ccos2pi(11, 16),
                               (float)ccos2pi(12, 16), (float)ccos2pi(13, 16), (float)ccos2pi(14, 16),
                               (float)ccos2pi(15, 16) };
    constexpr float sq[16] = { (float)csin2pi(0, 16),  (float)csin2pi(1, 16),  (float)csin2pi(2, 16),
                               (float)csin2pi(3, 16),  (float)csin2pi(4, 16),  (float)csin2pi(5, 16),
                               (float)csin2pi(6, 16),  (float)csin2pi(7, 16),  (float)csin2pi(8, 16),
                               (float)csin2pi(9, 16),  (float)csin2pi(10, 16), (float)csin2pi(11, 16),
                               (float)csin2pi(12, 16), (float)csin2pi(13, 16), (float)csin2pi(14, 16),
                               (float)csin2pi(15, 16) };
    const int l      = __lane_id();
    const int i      = l & 31;
    const int h      = l >> 5;
    const int m      = i & 15;
    const bool imrow = i >= 16;
    mfma16 K;
#pragma unroll
    for (int s = 0; s < 8; ++s)
    {
        const int q    = (m * (8 * h + s)) & 15;
        const float c  = cq[q];
        const float sg = INV ? -sq[q] : sq[q]; // forward F = cos - i sin
        // row i < 16: Re y_m = sum c Re x + sg Im x ; row i >= 16: Im y_m = sum -sg Re x + c Im x
        const float a0 = imrow ? -sg : c;
        const float a1 = imrow ? c : sg;
        __bf16 hi, lo;
        split_bf16(a0, hi, lo);
        K.a0h[s] = hi;
        K.a0l[s] = lo;
        split_bf16(a1, hi, lo);
        K.a1h[s] = hi;
        K.a1l[s] = lo;
    }
    return K;
}

// 16-point DFTs of the 32 butterflies whose inputs r in [8h, 8h+8) this lane holds in g
__device__ __forceinline__ f32x16 mfma16_group(const cx<float> (&g)[8], const mfma16& K)
{
    bf16x8 b0h, b0l, b1h, b1l;
#pragma unroll
    for (int s = 0; s < 8; ++s)
    {
        __bf16 hi, lo;
        split_bf16(g[s].re, hi, lo);
        b0h[s] = hi;
        b0l[s] = lo;
        split_bf16(g[s].im, hi, lo);
        b1h[s] = hi;
        b1l[s] = lo;
    }
    f32x16 acc = {};
    acc        = __builtin_amdgcn_mfma_f32_32x32x16_bf16(K.a0h, b0h, acc, 0, 0, 0);
    acc        = __builtin_amdgcn_mfma_f32_32x32x16_bf16(K.a0h, b0l, acc, 0, 0, 0);
    acc        = __builtin_amdgcn_mfma_f32_32x32x16_bf16(K.a0l, b0h, acc, 0, 0, 0);
    acc        = __builtin_amdgcn_mfma_f32_32x32x16_bf16(K.a1h, b1h, acc, 0, 0, 0);
    acc        = __builtin_amdgcn_mfma_f32_32x32x16_bf16(K.a1h, b1l, acc, 0, 0, 0);
    acc        = __builtin_amdgcn_mfma_f32_32x32x16_bf16(K.a1l, b1h, acc, 0, 0, 0);
    return acc;
}

// v_permlane32_swap on the pair (a, b): the upper half-wave of a and the lower half-wave of b trade
// places. Lane l < 32 keeps its a and receives the a of lane l + 32 in b; lane l >= 32 keeps its b
// and receives the b of lane l - 32 in a. One instruction moves both registers.
__device__ __forceinline__ void swap32(float& a, float& b)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a            = __uint_as_float(r[0]);
    b            = __uint_as_float(r[1]);
}

// Wave-cooperative radix-16 butterfly: every lane of the wave holds the 16 inputs of its own
// butterfly in u and gets its 16 outputs back. Swapping the pair (u[s], u[8+s]) leaves lane l < 32
// with input s of butterflies l and l + 32, and lane l >= 32 with input 8 + s of butterflies l - 32 and
// l: inputs [8h, 8h+8) of butterflies l%32 (group 0) and 32 + l%32 (group 1), the B operand layout.
// Both groups run on MFMA; lane h then holds outputs m(h, q) of both groups, and swapping each pair
// (group 0, group 1) returns outputs m(0, q) and m(1, q) of its own butterfly to every lane.
__device__ __forceinline__ void bfly16_mfma(cx<float> (&u)[16], const mfma16& K)
{
#pragma unroll
    for (int s = 0; s < 8; ++s)
    {
        swap32(u[s].re, u[8 + s].re);
        swap32(u[s].im, u[8 + s].im);
    }
    cx<float> g0[8], g1[8];
#pragma unroll
    for (int s = 0; s < 8; ++s)
    {
        g0[s] = u[s];
        g1[s] = u[8 + s];
    }
    const f32x16 d0 = mfma16_group(g0, K);
    const f32x16 d1 = mfma16_group(g1, K);
    // slot q holds output m = 4h + q (q < 4) or 8 + 4h + q - 4 (q >= 4): Re in d[q], Im in d[8 + q]
#pragma unroll
    for (int q = 0; q < 8; ++q)
    {
        cx<float> a = { d0[q], d0[8 + q] };
        cx<float> b = { d1[q], d1[8 + q] };
        swap32(a.re, b.re);
        swap32(a.im, b.im);
        u[q < 4 ? q : 4 + q]     = a; // output m(0, q)
        u[q < 4 ? 4 + q : 8 + q] = b; // output m(1, q)
    }
}

// Workgroup barrier for the LDS exchanges: the fences cover the local address space only, so it
// does not wait for outstanding global memory operations.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// the thread whose positions feed the second butterfly of thread t in a mirrored last pass
template <int MODE, int TP>
__device__ __forceinline__ int mirror_lane(int t)
{
    return MODE == 1 ? TP - 1 - t : (TP - t) & (TP - 1);
}

// pass P of the schedule on the registers v (strided set in, strided set out after the exchange).
// MIRROR0 (first pass of an even number M of butterflies per thread): butterfly m >= M/2 of thread t
// is TPT-1-t + m TPT instead of t + m TPT, so that x[q] and x[N-1-q] are inputs of the same thread's
// butterflies m and M-1-m (the mirrored input of the one-kernel DCT-II then needs no exchange).
// MIRRORL (last pass of two butterflies per thread): slot m + 2r holds X[j_m + 2r TPT] with j_0 = t and
// j_1 = TPT + mirror(t), the exchange before it reading the odd slots from the positions of thread
// mirror(t). mirror_rev (mirror(t) = TPT-1-t) gives one thread both X[q] and X[N-1-q]; mirror_neg
// (mirror(t) = (TPT-t) mod TPT) both X[q] and X[N-q] (thread 0's butterflies 0 and TPT pair with
// themselves).
constexpr int mirror_none = 0, mirror_rev = 1, mirror_neg = 2;
template <typename T, int L, bool INV, int ROLE, int PREC, int P, bool MIRROR0 = false, int MIRRORL = mirror_none>
__device__ __forceinline__ void run_pass(regs<T, L, ROLE, PREC>& v, cx<T>* lds, int tb, int t, const cx<T>* tw,
                                         const mfma16& K)
{
    using G          = geom<T, L, ROLE, PREC>;
    using S          = typename G::S;
    constexpr int E  = S::E;
    constexpr int TP = S::TPT;
    constexpr int R  = 1 << S::lr(P);
    constexpr int NS = 1 << S::lns(P);
    constexpr int M  = E / R;
    // v[m + r*M] holds x[j + r*N/R] for butterfly j = t + m*TPT
    if constexpr (NS > 1)
    {
#pragma unroll
        for (int m = 0; m < M; ++m)
        {
            const int j  = MIRRORL != mirror_none && P + 1 == S::NP && m >= M / 2
                               ? mirror_lane<MIRRORL, TP>(t) + m * TP
                               : t + m * TP;
            const int kt = j & (NS - 1);
            cx<T> w[R];
            if constexpr (R >= 8)
            {
                // W^(kt*r) from the row's entries at r = 2^b, one complex product per value: fewer
                // loads than the whole row (measured faster everywhere once the one-workgroup and row
                // kernels are compiled for the waves their LDS allows: the f32 2^14 and 2^15
                // one-workgroup kernels and the f64 2^10 row kernel had taken a wave per SIMD more
                // registers and kept whole-row loads, 5-14% slower)
                const cx<T>* row = tw + S::twoff(P) + kt * R;
                w[0]             = { T(1), T(0) };
#pragma unroll
                for (int b = 0; (1 << b) < R; ++b)
                    w[1 << b] = row[1 << b];
#pragma unroll
                for (int r = 3; r < R; ++r)
                    if ((r & (r - 1)) != 0)
                        w[r] = cmul(w[r - (1 << (ilog2c(r + 1) - 1))], w[1 << (ilog2c(r + 1) - 1)]);
            }
            else
                load_table<T, R>(tw + S::twoff(P) + kt * R, w);
#pragma unroll
            for (int r = 1; r < R; ++r)
                v[m + r * M] = ctw<INV>(v[m + r * M], w[r]);
        }
    }
#pragma unroll
    for (int m = 0; m < M; ++m)
    {
        cx<T> u[R];
#pragma unroll
        for (int r = 0; r < R; ++r)
            u[r] = v[m + r * M];
        if constexpr (PREC != prec_full && R == 16)
            bfly16_mfma(u, K);
        else
            butterfly<R, INV>(u);
#pragma unroll
        for (int r = 0; r < R; ++r)
            v[m + r * M] = u[r];
    }
    if constexpr (P + 1 < S::NP)
    {
        // output r of butterfly j goes to (j - j mod Ns)*R + j mod Ns + r*Ns. With H rounds, round h
        // moves positions [h*N/H, (h+1)*N/H): on a pass before the last these are exactly the outputs
        // of butterflies m in [h*M/H, (h+1)*M/H) (the top bits of the position are those of j / Ns)
        // and the inputs k in [h*E/H, (h+1)*E/H) of the next pass.
        // Round h reads into the registers it has just written out (slots m + r*M of its butterflies),
        // so the values of later rounds are never overwritten; a compile-time permutation then
        // restores the order v[k] = x[t + k*TPT].
        constexpr int H  = G::H;
        constexpr int MH = M / H;
#pragma unroll
        for (int h = 0; h < H; ++h)
        {
#pragma unroll
            for (int m = h * MH; m < (h + 1) * MH; ++m)
            {
                const int j    = MIRROR0 && m >= M / 2 ? TP - 1 - t + m * TP : t + m * TP;
                const int kt   = j & (NS - 1);
                const int base = (j - kt) * R + kt - h * G::NH;
#pragma unroll
                for (int r = 0; r < R; ++r)
                    lds[laddr<ROLE, G::W, G::LD>(tb, base + r * NS)] = v[m + r * M];
            }
            lds_barrier();
#pragma unroll
            for (int i = 0; i < E / H; ++i)
            {
                // position t + k TPT feeds slot k of the next pass; a mirrored last pass takes its odd
                // slots (its second butterfly) from position mirror(t) + k TPT
                const int k  = h * (E / H) + i;
                const int tk = MIRRORL != mirror_none && P + 2 == S::NP && (k & 1) ? mirror_lane<MIRRORL, TP>(t) : t;
                v[h * MH + i % MH + (i / MH) * M] = lds[laddr<ROLE, G::W, G::LD>(tb, tk + k * TP - h * G::NH)];
            }
            lds_barrier();
        }
        if constexpr (H > 1)
        {
            cx<T> u[E];
#pragma unroll
            for (int h = 0; h < H; ++h)
#pragma unroll
                for (int i = 0; i < E / H; ++i)
                    u[h * (E / H) + i] = v[h * MH + i % MH + (i / MH) * M];
#pragma unroll
            for (int k = 0; k < E; ++k)
                v[k] = u[k];
        }
        run_pass<T, L, INV, ROLE, PREC, P + 1, false, MIRRORL>(v, lds, tb, t, tw, K);
    }
    static_assert(!MIRROR0 || (P == 0 && NS == 1 && M % 2 == 0 && P + 1 < S::NP),
                  "mirrored first pass: an even number of butterflies");
    static_assert(MIRRORL == mirror_none || (S::NP >= 2 && E / (1 << S::lr(S::NP - 1)) == 2),
                  "mirrored last pass: two butterflies");
}

// power-of-two scale 2^k of a value array whose largest |re|, |im| is mx: the scaled values stay below
// 2^15, inside binary16's range
__device__ __forceinline__ int mid_scale_exp(float mx)
{
    int e = 0;
    (void)frexpf(mx, &e); // mx < 2^e
    return mx > 0.f ? 15 - e : 0;
}

// the scale exponent of all values the workgroup holds (every thread calls it): wave maxima, then one
// value per wave through LDS; the LDS is free again when it returns
template <typename T, int E, int BS>
__device__ __forceinline__ int workgroup_scale(const cx<T> (&v)[E], cx<T>* lds)
{
    float mx = 0.f;
#pragma unroll
    for (int k = 0; k < E; ++k)
        mx = fmaxf(mx, fmaxf(fabsf((float)v[k].re), fabsf((float)v[k].im)));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        mx = fmaxf(mx, __shfl_xor(mx, off));
    float* red = reinterpret_cast<float*>(lds);
    lds_barrier();
    if (__lane_id() == 0)
        red[threadIdx.x >> 6] = mx;
    lds_barrier();
    float bm = 0.f;
#pragma unroll
    for (int w = 0; w < (BS + 63) / 64; ++w)
        bm = fmaxf(bm, red[w]);
    lds_barrier();
    return mid_scale_exp(bm);
}

template <typename T, int L, bool INV, int ROLE, int PREC>
__global__ void __launch_bounds__((geom<T, L, ROLE, PREC>::BS),
                                  (ROLE == role_direct || ROLE == role_row ? geom<T, L, ROLE, PREC>::lds_waves : 1))
    k_pow2(args<T> a)
{
    using G          = geom<T, L, ROLE, PREC>;
    using S          = typename G::S;
    constexpr int N  = S::N;
    constexpr int E  = S::E;
    constexpr int TP = S::TPT;
    constexpr int W  = G::W;
    constexpr int BS = G::BS;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    mfma16 K{};
    if constexpr (PREC != prec_full)
        K = make_mfma16<INV>();

    unsigned blk = blockIdx.x;
    if constexpr (ROLE != role_direct)
    {
        // workgroup b runs on XCD b mod 8: give each XCD a contiguous range of column/row groups
        const unsigned per = gridDim.x / 8;
        blk                = (blockIdx.x % 8) * per + blockIdx.x / 8;
        if (blk >= a.nblk)
            return; // whole workgroup, before any barrier
    }
    const int tid     = threadIdx.x;
    constexpr bool COLS = ROLE == role_col || ROLE == role_pass; // lanes on consecutive columns
    const int tb      = COLS ? tid % W : tid / TP;
    const int t       = COLS ? tid / W : tid % TP;
    const long long g = (long long)blk * W + tb;
    const bool valid  = g < a.count;
    cx<T> v[E];
    // stores that fill whole 128-byte lines go out nontemporal (measured: the one-workgroup f32 kernels
    // below 2^15 points ran up to 5% slower with them, so they keep ordinary stores)
    constexpr bool NT = ROLE == role_direct ? sizeof(T) == 8 || L >= 15
                                            : ROLE == role_pass || W * (int)sizeof(cx<T>) >= 128;

    if constexpr (ROLE == role_direct && G::staged)
    {
        const long long base = (long long)blk * W * N;
        const long long lim  = a.count * N - base;
#pragma unroll
        for (int k = 0; k < E; ++k)
        {
            const int e = tid + k * BS;
            v[k]        = e < lim ? ld_elem<T, PREC>(a.in, base + e) : cx<T>{ T(0), T(0) };
        }
#pragma unroll
        for (int k = 0; k < E; ++k)
        {
            const int e       = tid + k * BS;
            lds[e + (e >> 4)] = v[k];
        }
        lds_barrier();
#pragma unroll
        for (int k = 0; k < E; ++k)
            v[k] = lds[laddr<ROLE, W, G::LD>(tb, t + k * TP)];
        lds_barrier();
    }
    else if constexpr (ROLE == role_col || ROLE == role_pass)
    {
        const long long bl   = g >> a.lo;
        const long long col  = g & ((1ll << a.lo) - 1);
        const long long base = (bl << (a.lo + L)) + col;
        bool scaled_in       = false;
        if constexpr (ROLE == role_pass && PREC != prec_full)
            scaled_in = a.in_scales != nullptr;
        if (scaled_in)
        {
            // binary16 intermediate of the previous pass: element o was written by its butterfly
            // j' = (o >> (ns' + l')) << ns' | (o mod 2^ns'), i.e. by workgroup (bl 2^(Ltot - l') + j') >> wlog'
            const long long nsm = (1ll << a.in_ns) - 1;
#pragma unroll
            for (int k = 0; k < E; ++k)
            {
                v[k] = { T(0), T(0) };
                if (valid)
                {
                    const long long o  = col + ((long long)(t + k * TP) << a.lo);
                    const long long jp = ((o >> (a.in_ns + a.in_l)) << a.in_ns) | (o & nsm);
                    const long long gp = (bl << (a.lo + L - a.in_l)) + jp;
                    const int sc       = a.in_scales[gp >> a.in_wlog];
                    const cx<T> h      = ld_elem<T, prec_f16>(a.in, base + ((long long)(t + k * TP) << a.lo));
                    v[k]               = { ldexpf(h.re, -sc), ldexpf(h.im, -sc) };
                }
            }
        }
        else
        {
#pragma unroll
            for (int k = 0; k < E; ++k)
                v[k] = valid ? ld_elem<T, PREC>(a.in, base + ((long long)(t + k * TP) << a.lo)) : cx<T>{ T(0), T(0) };
        }
        if constexpr (ROLE == role_pass)
        {
            if (a.ns_log > 0)
            {
                // input r = t + k*TPT of butterfly j is multiplied by W_{Ns*R}^(r * kt), kt = j mod Ns:
                // W^(kt*t) times W^(kt*TPT*k), the latter built from the table values W^(kt*TPT*2^b)
                // with one complex product per value (rounding depth at most log2(E) + 1)
                const unsigned long long kt   = (unsigned long long)col & ((1ull << a.ns_log) - 1);
                const unsigned long long mask = (1ull << a.lbits) - 1;
                const unsigned long long nsr  = (1ull << (a.ns_log + L)) - 1; // Ns*R - 1
                auto table = [&](unsigned long long m) { return cmul(a.ta[m >> a.lbits], a.tb[m & mask]); };
                constexpr int LEE = ilog2c(E);
                cx<T> pw[LEE > 0 ? LEE : 1];
#pragma unroll
                for (int b = 0; b < LEE; ++b)
                    pw[b] = table((kt * (unsigned long long)(TP << b)) & nsr);
                cx<T> w[E];
                w[0] = table(kt * (unsigned long long)t);
#pragma unroll
                for (int k = 1; k < E; ++k)
                {
                    const int hb = ilog2c(k + 1) - 1; // top bit of k
                    w[k]         = cmul(w[k - (1 << hb)], pw[hb]);
                }
#pragma unroll
                for (int k = 0; k < E; ++k)
                    v[k] = ctw<INV>(v[k], w[k]);
            }
        }
    }
    else if constexpr (ROLE == role_row && PREC != prec_full)
    {
        // binary16 intermediate: element (col = i) carries the scale of its column-kernel workgroup
        const long long base = g * N;
#pragma unroll
        for (int k = 0; k < E; ++k)
        {
            v[k] = { T(0), T(0) };
            if (valid)
            {
                const long long e  = base + t + k * TP;
                const long long bl = g >> a.lo;
                const long long gc = (bl << L) + t + k * TP; // column (bl, col) over the batch
                const int sc      = a.scales[gc >> a.gshift];
                const cx<T> h     = ld_elem<T, prec_f16>(a.in, e);
                v[k]              = { ldexpf(h.re, -sc), ldexpf(h.im, -sc) };
            }
        }
    }
    else
    {
        const long long base = g * N + t;
#pragma unroll
        for (int k = 0; k < E; ++k)
            v[k] = valid ? ld_elem<T, ROLE == role_row ? mid_prec<PREC> : PREC>(a.in, base + k * TP)
                         : cx<T>{ T(0), T(0) };
    }

    run_pass<T, L, INV, ROLE, PREC, 0>(v, lds, tb, t, a.tw, K);
    // v[k] = X[t + k*TPT]

    if constexpr (ROLE == role_direct && G::staged)
    {
#pragma unroll
        for (int k = 0; k < E; ++k)
            lds[laddr<ROLE, W, G::LD>(tb, t + k * TP)] = v[k];
        lds_barrier();
        const long long base = (long long)blk * W * N;
        const long long lim  = a.count * N - base;
#pragma unroll
        for (int k = 0; k < E; ++k)
        {
            const int e = tid + k * BS;
            if (e < lim)
                st_elem<T, PREC, NT>(a.out, base + e, lds[e + (e >> 4)]);
        }
    }
    else if constexpr (ROLE == role_direct)
    {
        if (valid)
        {
            const long long base = g * N + t;
#pragma unroll
            for (int k = 0; k < E; ++k)
                st_elem<T, PREC, NT>(a.out, base + k * TP, v[k]);
        }
    }
    else if constexpr (ROLE == role_pass)
    {
        const long long bl  = g >> a.lo;
        const long long col = g & ((1ll << a.lo) - 1);
        // half precision: an intermediate is written as binary16 scaled by 2^sc, one sc per workgroup
        int sc = 0;
        bool mid = false;
        if constexpr (PREC != prec_full)
        {
            mid = a.out_mid != 0;
            if (mid)
            {
                sc = workgroup_scale<T, E, BS>(v, lds);
                if (tid == 0)
                    a.scales[blk] = sc;
            }
        }
        auto put = [&](long long idx, cx<T> x)
        {
            if constexpr (PREC != prec_full)
            {
                if (mid)
                {
                    st_elem<T, prec_f16>(a.out, idx, { ldexpf(x.re, sc), ldexpf(x.im, sc) });
                    return;
                }
            }
            st_elem<T, PREC, NT>(a.out, idx, x);
        };
        if (a.ns_log == 0)
        {
            // first pass: output r of butterfly j goes to j*R + r, so the workgroup's W butterflies
            // fill one contiguous run of W*R values; write it through LDS in order (with H rounds,
            // round h moves outputs [h*N/H, (h+1)*N/H) of every butterfly, runs of N/H values)
            constexpr int H     = G::H;
            constexpr int NH    = G::NH;
            const long long g0  = (long long)blk * W;
            const long long b0  = g0 >> a.lo;
            const long long run = (b0 << (a.lo + L)) + ((g0 & ((1ll << a.lo) - 1)) << L);
            const long long lim = (a.count - g0) << L;
#pragma unroll
            for (int h = 0; h < H; ++h)
            {
                if (h > 0)
                    lds_barrier();
#pragma unroll
                for (int k = h * (E / H); k < (h + 1) * (E / H); ++k)
                    lds[laddr<ROLE, W, G::LD>(tb, t + k * TP - h * NH)] = v[k];
                lds_barrier();
#pragma unroll
                for (int k = 0; k < E / H; ++k)
                {
                    const int q         = tid + k * BS; // < W * NH
                    const int tq        = q / NH;
                    const int iq        = q - tq * NH;
                    const long long o   = (long long)tq * N + h * NH + iq;
                    if (o < lim)
                        put(run + o, lds[laddr<ROLE, W, G::LD>(tq, iq)]);
                }
            }
        }
        else if (valid)
        {
            // output r of butterfly j goes to (j - j mod Ns)*R + j mod Ns + r*Ns
            const long long kt   = col & ((1ll << a.ns_log) - 1);
            const long long base = (bl << (a.lo + L)) + ((col >> a.ns_log) << (a.ns_log + L)) + kt;
#pragma unroll
            for (int k = 0; k < E; ++k)
                put(base + ((long long)(t + k * TP) << a.ns_log), v[k]);
        }
    }
    else if constexpr (ROLE == role_col)
    {
        const long long bl            = g >> a.lo;
        const long long col           = g & ((1ll << a.lo) - 1);
        const long long base          = (bl << (a.lo + L)) + col;
        const unsigned long long mask = (1ull << a.lbits) - 1;
        if (a.ta) // four-step twiddles (absent when transforming one axis of a multidimensional array)
        {
            // W_n^(col*k2), k2 = t + k*TPT: W^(col*t) times W^(col*TPT*k), the latter built from the
            // table values W^(col*TPT*2^b) with one complex product per value
            const unsigned long long nm = (1ull << (a.lo + L)) - 1; // n - 1
            auto table = [&](unsigned long long m) { return cmul(a.ta[m >> a.lbits], a.tb[m & mask]); };
            constexpr int LEE = ilog2c(E);
            cx<T> pw[LEE > 0 ? LEE : 1];
#pragma unroll
            for (int b = 0; b < LEE; ++b)
                pw[b] = table(((unsigned long long)col * (unsigned long long)(TP << b)) & nm);
            cx<T> w[E];
            w[0] = table((unsigned long long)col * (unsigned long long)t);
#pragma unroll
            for (int k = 1; k < E; ++k)
            {
                const int hb = ilog2c(k + 1) - 1; // top bit of k
                w[k]         = cmul(w[k - (1 << hb)], pw[hb]);
            }
#pragma unroll
            for (int k = 0; k < E; ++k)
                v[k] = ctw<INV>(v[k], w[k]);
        }
        if constexpr (PREC == prec_full)
        {
            if (valid)
            {
#pragma unroll
                for (int k = 0; k < E; ++k)
                    st_elem<T, PREC, NT>(a.out, base + ((long long)(t + k * TP) << a.lo), v[k]);
            }
        }
        else
        {
            // one scale for the workgroup's columns: wave maxima through LDS (free after the passes)
            float mx = 0.f;
#pragma unroll
            for (int k = 0; k < E; ++k)
                mx = fmaxf(mx, fmaxf(fabsf((float)v[k].re), fabsf((float)v[k].im)));
#pragma unroll
            for (int off = 32; off > 0; off >>= 1)
                mx = fmaxf(mx, __shfl_xor(mx, off));
            float* red = reinterpret_cast<float*>(lds);
            if (__lane_id() == 0)
                red[tid >> 6] = mx;
            lds_barrier();
            float bm = 0.f;
#pragma unroll
            for (int w = 0; w < BS / 64; ++w)
                bm = fmaxf(bm, red[w]);
            const int sc = mid_scale_exp(bm);
            if (tid == 0)
                a.scales[blk] = sc;
            if (valid)
            {
#pragma unroll
                for (int k = 0; k < E; ++k)
                    st_elem<T, prec_f16>(a.out, base + ((long long)(t + k * TP) << a.lo),
                                         { ldexpf(v[k].re, sc), ldexpf(v[k].im, sc) });
            }
        }
    }
    else
    {
        // transposed store: row index fastest, so lanes write W consecutive outputs k2 + n2*k1; with
        // H rounds, round h moves rows k1 in [h*N/H, (h+1)*N/H) (values k in [h*E/H, (h+1)*E/H))
        constexpr int H    = G::H;
        const int tb2      = tid % W;
        const int i0       = tid / W;
        const long long g2 = (long long)blk * W + tb2;
        const long long bl   = g2 >> a.lo;
        const long long k2   = g2 & ((1ll << a.lo) - 1);
        const long long base = (bl << (a.lo + L)) + k2;
#pragma unroll
        for (int h = 0; h < H; ++h)
        {
            if (h > 0)
                lds_barrier();
#pragma unroll
            for (int k = h * (E / H); k < (h + 1) * (E / H); ++k)
                lds[laddr<ROLE, W, G::LD>(tb, t + k * TP - h * G::NH)] = v[k];
            lds_barrier();
            if (g2 < a.count)
            {
                // one running pointer: the outputs of a thread are TPT << lo apart
                const long long step = (long long)TP << a.lo;
                long long o          = base + ((long long)(i0 + h * (E / H) * TP) << a.lo);
#pragma unroll
                for (int k = h * (E / H); k < (h + 1) * (E / H); ++k)
                {
                    const int i = i0 + k * TP;
                    st_elem<T, PREC, NT>(a.out, o, lds[laddr<ROLE, W, G::LD>(tb2, i - h * G::NH)]);
                    o += step;
                }
            }
        }
    }
}

template <typename T, int L, bool INV, int ROLE, int PREC>
void launch(args<T> a, long long count, hipStream_t s)
{
    using G                  = geom<T, L, ROLE, PREC>;
    static const bool lds_ok = (allow_lds(k_pow2<T, L, INV, ROLE, PREC>, G::lds_bytes), true);
    (void)lds_ok;
    if (count <= 0)
        return;
    a.count             = count;
    const long long nb  = (count + G::W - 1) / G::W;
    const long long grd = ROLE == role_direct ? nb : (nb + 7) / 8 * 8;
    if (grd > 0x7fffffffll)
        throw std::runtime_error("batch too large for one launch");
    a.nblk            = (unsigned)nb;
    const long long es = PREC == prec_full ? (long long)sizeof(cx<T>) : 4;
    timing::span ts([] { return timing::kernel_name("k_pow2", timing::type_name<T>(), L, INV, ROLE, PREC); },
                    2 * es * (count << L), s);
    hipLaunchKernelGGL((k_pow2<T, L, INV, ROLE, PREC>), dim3((unsigned)grd), dim3(G::BS), G::lds_bytes, s, a);
    hip_check(hipGetLastError(), "power-of-two FFT launch");
}

template <typename T, bool INV, int ROLE, int PREC, int LO, int HI>
void launch_len(int L, const args<T>& a, long long count, hipStream_t s)
{
    if constexpr (LO <= HI)
    {
        if (L == LO)
            return launch<T, LO, INV, ROLE, PREC>(a, count, s);
        return launch_len<T, INV, ROLE, PREC, LO + 1, HI>(L, a, count, s);
    }
    else
        throw std::logic_error("no power-of-two kernel for this length");
}

template <typename T, int ROLE, int PREC, int LO, int HI>
void launch_any(bool inv, int L, const args<T>& a, long long count, hipStream_t s)
{
    if (inv)
        launch_len<T, true, ROLE, PREC, LO, HI>(L, a, count, s);
    else
        launch_len<T, false, ROLE, PREC, LO, HI>(L, a, count, s);
}

// columns per workgroup of the column kernel of 2^L points at precision PREC (the row kernel's scale
// grouping; the full-precision and half-precision column kernels differ in width)
template <typename T, int PREC, int LO, int HI>
int col_width(int L)
{
    if constexpr (LO <= HI)
    {
        if (L == LO)
            return geom<T, LO, role_col, PREC>::W;
        return col_width<T, PREC, LO + 1, HI>(L);
    }
    else
        return 1;
}

// butterflies per workgroup of the multi-pass kernel of radix 2^L
template <typename T, int PREC, int LO, int HI>
int pass_width(int L)
{
    if constexpr (LO <= HI)
    {
        if (L == LO)
            return geom<T, LO, role_pass, PREC>::W;
        return pass_width<T, PREC, LO + 1, HI>(L);
    }
    else
        return 1;
}

// threads per transform of the one-workgroup kernel of 2^L points
template <typename T, int LO, int HI>
int direct_threads(int L)
{
    if constexpr (LO <= HI)
    {
        if (L == LO)
            return geom<T, LO, role_direct>::S::TPT;
        return direct_threads<T, LO + 1, HI>(L);
    }
    else
        return 1;
}

// pass tables of 2^L points (layout of pass_twoff), FP64-generated
template <typename T>
void gen_pass_tables(int L, cx<T>* tw, hipStream_t s)
{
    const int le = log_e<T>;
    for (int p = 0; p < pass_count(L, le); ++p)
    {
        const long long ns = 1ll << pass_lns(L, le, p);
        const int r        = 1 << pass_lr(L, le, p);
        if (ns > 1)
            gen_twiddles(tw + pass_twoff(L, le, p), ns * r, ns * r, 0, 0, tw_radix_full, r, s);
    }
}

inline int log2_exact(size_t n)
{
    if (n == 0 || (n & (n - 1)))
        return -1;
    int l = 0;
    while ((size_t(1) << l) < n)
        ++l;
    return l;
}

// ---- DCT-II in one kernel (dct_plan, include/kfr/dft/fft.hpp:927-970, in Makhoul's real-transform
// form): for n = 2N reals, z[q] = (x[4q], x[4q+2]) and z[N-1-q] = (x[4q+3], x[4q+1]) (q < N/2) is the
// mirrored input as N complex values; Z = DFT_N(z); V[k] = A[k] + W_n^k B[k] with A, B the spectra of
// the even and odd samples (from Z[k] and Z[N-k]); y[k] = Re(V[k] W_4n^k), y[n-k] = -Im(V[k] W_4n^k),
// y[N] = Re((Re Z0 - Im Z0) W_4n^N). Thread t holds z[t + k TPT]: slots k < E/2 take z[q] straight
// from one coalesced load of x[4q..4q+3], whose other half z[N-1-q] is slot E-1-k of thread TPT-1-t and
// crosses through LDS, unless the first pass has several butterflies per thread: they are then
// assigned in mirror pairs and each thread loads all of its inputs. After the passes, with two
// butterflies per thread in the last pass, those are assigned so that one thread holds Z[i] and Z[N-i]
// and writes the outputs of both; otherwise one LDS round of N/2 values brings Z[N-i] to the slot
// i < N/2, which writes the outputs of both i and N-i.
template <typename T>
struct alignas(4 * sizeof(T)) quad
{
    T a, b, c, d;
};

// a quad store, nontemporal if NT (the one-kernel DCT-III: measured 2.5-9% faster with it on 1 GiB
// batches; the one-kernel DCT-II, R2C and C2R measured no better or slower and store normally)
template <bool NT, typename T>
__device__ __forceinline__ void put(quad<T>* p, quad<T> v)
{
    using V = T __attribute__((ext_vector_type(4)));
    if constexpr (NT)
        __builtin_nontemporal_store(V{ v.a, v.b, v.c, v.d }, reinterpret_cast<V*>(p));
    else
        *p = v;
}

template <typename T, int L>
__global__ void __launch_bounds__((geom<T, L, role_direct>::BS), (geom<T, L, role_direct>::lds_waves))
k_dct2(const T* __restrict__ x, T* __restrict__ y, const cx<T>* __restrict__ tw, const cx<T>* __restrict__ wn,
       const cx<T>* __restrict__ w4n, long long count)
{
    using G          = geom<T, L, role_direct>;
    using S          = typename G::S;
    constexpr int N  = S::N;
    constexpr int E  = S::E;
    constexpr int TP = S::TPT;
    constexpr int W  = G::W;
    constexpr int HN = N / 2;
    static_assert(!G::staged && E % 2 == 0 && G::H <= 2, "fused DCT needs whole-thread transforms");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const mfma16 K{};
    const int tid     = threadIdx.x;
    const int tb      = tid / TP;
    const int t       = tid % TP;
    const long long g = (long long)blockIdx.x * W + tb;
    const bool valid  = g < count;
    const long long n = 2ll * N;
    auto at           = [&](int i) -> cx<T>& { return lds[laddr<role_direct, W, G::LD>(tb, i)]; };
    cx<T> v[E];
    constexpr int M0 = E >> S::lr(0); // butterflies per thread in the first pass
    // two butterflies per thread in the last pass: assigned in pairs whose outputs are Z[i] and Z[N-i]
    constexpr int ML = G::EX == 2 && S::NP >= 2 ? mirror_neg : mirror_none;
    if constexpr (M0 >= 2 && S::NP > 1)
    {
        // The first pass's butterflies of thread t are j_m = t + m TPT for m < M0/2 and
        // j_m = TPT-1-t + m TPT for m >= M0/2, so j_m + j_(M0-1-m) = N/R0 - 1 (inputs j + r N/R0 in
        // slot m + r M0). For r < R0/2 quad q = j_m + r N/R0 holds z[q] (input r of j_m) and
        // z[N-1-q] (input R0-1-r of j_(M0-1-m)): the thread loads all of its inputs itself, no exchange
        constexpr int R0  = E / M0;
        const quad<T>* xq = reinterpret_cast<const quad<T>*>(x + g * n);
#pragma unroll
        for (int m = 0; m < M0; ++m)
        {
            const int jm = m < M0 / 2 ? t + m * TP : TP - 1 - t + m * TP;
#pragma unroll
            for (int r = 0; r < R0 / 2; ++r)
            {
                quad<T> q4 = { T(0), T(0), T(0), T(0) };
                if (valid)
                    q4 = xq[jm + r * (M0 * TP)];
                v[m + r * M0]                       = { q4.a, q4.c };
                v[(M0 - 1 - m) + (R0 - 1 - r) * M0] = { q4.d, q4.b };
            }
        }
        run_pass<T, L, false, role_direct, prec_full, 0, true, ML>(v, lds, tb, fresh(t), tw, K);
    }
    else
    {
#pragma unroll
        for (int k = 0; k < E / 2; ++k)
        {
            quad<T> q4 = { T(0), T(0), T(0), T(0) };
            if (valid)
                q4 = reinterpret_cast<const quad<T>*>(x + g * n)[t + k * TP];
            v[k]                                    = { q4.a, q4.c };
            at((E / 2 - 1 - k) * TP + (TP - 1 - t)) = { q4.d, q4.b }; // z[N-1-q] - N/2
        }
        lds_barrier();
#pragma unroll
        for (int k = E / 2; k < E; ++k)
            v[k] = at((k - E / 2) * TP + t);
        lds_barrier();
        run_pass<T, L, false, role_direct, prec_full, 0, false, ML>(v, lds, tb, fresh(t), tw, K);
    }
    // v[k] = Z[t + k TPT]. Output pair i, n - i needs Z[i] and Z[N - i] (Z[0] for i = 0).
    T* yb     = y + g * n;
    // i = 0: V[0] = Re Z0 + Im Z0 and V[N] = Re Z0 - Im Z0, both real
    auto emit0 = [&](cx<T> a)
    {
        yb[0]          = a.re + a.im;
        const cx<T> wN = w4n[N];
        yb[N]          = (a.re - a.im) * wN.re;
    };
    // 0 < i < N
    auto emit = [&](int i, cx<T> a, cx<T> c)
    {
        // A = (Z[i] + conj Z[N-i]) / 2, B = (Z[i] - conj Z[N-i]) / 2i, V = A + W_n^i B
        const cx<T> A  = { T(0.5) * (a.re + c.re), T(0.5) * (a.im - c.im) };
        const cx<T> B  = { T(0.5) * (a.im + c.im), -T(0.5) * (a.re - c.re) };
        const cx<T> wi = wn[i];
        const cx<T> V  = { A.re + wi.re * B.re - wi.im * B.im, A.im + wi.re * B.im + wi.im * B.re };
        const cx<T> w  = w4n[i];
        yb[i]          = V.re * w.re - V.im * w.im;
        yb[n - i]      = -(V.re * w.im + V.im * w.re);
    };
    if constexpr (ML == mirror_neg)
    {
        // slot m + 2r holds Z[j_m + 2r TPT] with j_0 = t, j_1 = 2 TPT - t (thread 0: j_1 = TPT), so
        // Z[N-i] of i = t + 2r TPT is slot 1 + 2(R-1-r); thread 0's butterflies pair with themselves
        constexpr int R = E / 2;
        if (!valid)
            return;
        const int ts = fresh(t);
        if (ts != 0)
        {
#pragma unroll
            for (int r = 0; r < R; ++r)
            {
                const int i   = ts + 2 * r * TP;
                const cx<T> a = v[2 * r], c = v[1 + 2 * (R - 1 - r)];
                emit(i, a, c);
                emit(N - i, c, a);
            }
        }
        else
        {
            emit0(v[0]);
#pragma unroll
            for (int r = 1; r < R / 2; ++r)
            {
                const int i = 2 * r * TP;
                emit(i, v[2 * r], v[2 * (R - r)]);
                emit(N - i, v[2 * (R - r)], v[2 * r]);
            }
            emit(HN, v[R], v[R]);
#pragma unroll
            for (int r = 0; r < R / 2; ++r)
            {
                const int i = TP + 2 * r * TP;
                emit(i, v[1 + 2 * r], v[1 + 2 * (R - 1 - r)]);
                emit(N - i, v[1 + 2 * (R - 1 - r)], v[1 + 2 * r]);
            }
        }
        return;
    }
    else
    {
        // the upper slots publish Z; each lower slot i reads Z[N-i] and writes the outputs of both i
        // and N - i; slot N/2 pairs with itself. The indices come from a fresh copy of t, so none is
        // held through the passes.
        const int ts = fresh(t);
#pragma unroll
        for (int k = E / 2; k < E; ++k)
            at(ts + k * TP - HN) = v[k];
        lds_barrier();
        if (!valid)
            return;
#pragma unroll
        for (int k = 0; k < E / 2; ++k)
        {
            const int i = ts + k * TP;
            if (k == 0 && ts == 0)
                emit0(v[k]);
            else
            {
                const cx<T> c = at(N - i - HN);
                emit(i, v[k], c);
                emit(N - i, c, v[k]);
            }
        }
        if (ts == 0)
            emit(HN, v[E / 2], v[E / 2]);
    }
}

// DCT-III (dct_plan's inverse, unscaled) in one kernel, the mirror image of k_dct2: with V[k] = y'[k] W_4n^k
// (y'[0] = y[0]/2) and H[k] = (conj V[k] + V[n-k])/2, slot i forms
// Z'[i] = (H[i] + conj H[N-i]) + i (H[i] - conj H[N-i]) conj(W_n^i), the inverse N-point transform runs, and
// x[4q..4q+3] = (u[q].re, u[N-1-q].im, u[q].im, u[N-1-q].re) leaves as one 16-byte store per q < N/2.
// With two butterflies per thread in the last pass they are assigned in mirror pairs and each thread
// holds both u[q] and u[N-1-q]; otherwise u[N-1-q] crosses through LDS from slot E-1-k of thread
// TPT-1-t.
template <typename T, int L>
__global__ void __launch_bounds__((geom<T, L, role_direct>::BS), (geom<T, L, role_direct>::lds_waves))
k_dct3(const T* __restrict__ y, T* __restrict__ x, const cx<T>* __restrict__ tw, const cx<T>* __restrict__ wn,
       const cx<T>* __restrict__ w4n, long long count)
{
    using G          = geom<T, L, role_direct>;
    using S          = typename G::S;
    constexpr int N  = S::N;
    constexpr int E  = S::E;
    constexpr int TP = S::TPT;
    constexpr int W  = G::W;
    constexpr bool NT = !(sizeof(T) == 8 && L == 14); // nontemporal output (f64 2^14 measured 13% slower)
    constexpr int HN = N / 2;
    static_assert(!G::staged && E % 2 == 0 && G::H <= 2, "fused DCT needs whole-thread transforms");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const mfma16 K{};
    const int tid     = threadIdx.x;
    const int tb      = tid / TP;
    const int t       = tid % TP;
    const long long g = (long long)blockIdx.x * W + tb;
    const bool valid  = g < count;
    constexpr long long n = 2ll * N;
    auto at               = [&](int i) -> cx<T>& { return lds[laddr<role_direct, W, G::LD>(tb, i)]; };
    // a thread past the last transform loads the last one (its results are never stored), so the loads
    // need no branches
    const T* yb = y + (valid ? g : count - 1) * n;
    // H[kk] = (conj V[kk] + V[n-kk]) / 2 from a = y[kk], b = y[n-kk] and w = W_4n^kk, using
    // W_4n^(n-kk) = (-w.im, -w.re) (the table values swapped, so exactly the rounded table entry);
    // kk = 0 reads y[0] for both (V[n] = V[0]) and halves it
    auto Hs = [&](bool zero, T a, T b, cx<T> w) -> cx<T>
    {
        if (zero)
            return { T(0.5) * a * w.re, T(0) };
        const cx<T> w1 = { -w.im, -w.re };
        return { T(0.5) * (a * w.re + b * w1.re), T(0.5) * (b * w1.im - a * w.im) };
    };
    // slot i holds H[i] (two coalesced loads, one in reverse order); H[N] (the partner of slot 0) is
    // computed by every thread from y[N]
    // (a + conj c) + i (a - conj c) u with u = conj(W_n^i); slot N - i has u = -W_n^i, since
    // W_n^(N-i) = -conj(W_n^i), so one table value serves both slots of a pair
    auto pack = [&](cx<T> a, cx<T> c, cx<T> u) -> cx<T>
    {
        const cx<T> sm = { a.re + c.re, a.im - c.im };
        const cx<T> d  = { a.re - c.re, a.im + c.im };
        const cx<T> e  = cmul(d, u);
        return { sm.re - e.im, sm.im + e.re };
    };
    cx<T> v[E];
#pragma unroll
    for (int k = 0; k < E; ++k)
    {
        const int i = t + k * TP;
        v[k]        = Hs(i == 0, yb[i], yb[i == 0 ? 0 : n - i], w4n[i]);
        settle(v[k]);
    }
    const T yN     = yb[N];
    const cx<T> hN = Hs(false, yN, yN, w4n[N]);
    // the upper slots publish H; each lower slot i reads H[N-i], packs both slot i and slot N-i, and
    // leaves the latter in the position it read; the upper slots collect theirs after one barrier
#pragma unroll
    for (int k = E / 2; k < E; ++k)
        at(t + k * TP - HN) = v[k];
    lds_barrier();
    // the pack phase recomputes its offsets rather than keep the load phase's; slot 0 (k = 0, t = 0)
    // pairs with H[N] and slot N/2 (k = E/2, t = 0) with itself, selected without branches: slot 0
    // reads and rewrites position N/2 of the transform's LDS area, which nothing reads before the
    // passes overwrite it, and slot N/2 reads position 0
    const int tp = fresh(t);
#pragma unroll
    for (int k = 0; k < E / 2; ++k)
    {
        const int i     = tp + k * TP;
        const bool zero = k == 0 && tp == 0;
        cx<T>& slot     = at(N - i - HN);
        const cx<T> s0  = slot;
        const cx<T> hp  = zero ? hN : s0;
        const cx<T> w   = wn[i];
        slot            = pack(hp, v[k], { -w.re, -w.im });
        v[k]            = pack(v[k], hp, cconj(w));
        settle(v[k]);
    }
    lds_barrier();
#pragma unroll
    for (int k = E / 2; k < E; ++k)
    {
        const int i  = tp + k * TP;
        const bool h = k == E / 2 && tp == 0;
        const cx<T> c = at(i - HN);
        v[k]          = h ? pack(v[k], v[k], cconj(wn[HN])) : c;
    }
    lds_barrier();
    quad<T>* xq = reinterpret_cast<quad<T>*>(x + g * n);
    if constexpr (G::EX == 2 && S::NP >= 2)
    {
        // two butterflies per thread in the last pass, assigned in mirror pairs: slot m + 2r holds
        // u[j_m + 2r TPT] (j_0 = t, j_1 = 2 TPT-1-t), and u[N-1-q] of q = j_m + 2r TPT is output R-1-r
        // of the other butterfly, so every quad leaves from registers
        constexpr int R = E / 2;
        run_pass<T, L, true, role_direct, prec_full, 0, false, mirror_rev>(v, lds, tb, fresh(t), tw, K);
        if (!valid)
            return;
        const int ts = fresh(t);
#pragma unroll
        for (int r = 0; r < R / 2; ++r)
        {
            const cx<T> a0 = v[2 * r], a1 = v[1 + 2 * (R - 1 - r)];
            put<NT>(xq + ts + 2 * r * TP, quad<T>{ a0.re, a1.im, a0.im, a1.re });
            const cx<T> b0 = v[1 + 2 * r], b1 = v[2 * (R - 1 - r)];
            put<NT>(xq + (2 * TP - 1 - ts + 2 * r * TP), quad<T>{ b0.re, b1.im, b0.im, b1.re });
        }
    }
    else
    {
        run_pass<T, L, true, role_direct, prec_full, 0>(v, lds, tb, fresh(t), tw, K);
        // v[k] = u[t + k TPT]; the upper half of the slots crosses to the lower half in mirror order
        const int ts = fresh(t);
#pragma unroll
        for (int k = E / 2; k < E; ++k)
            at(ts + k * TP - HN) = v[k];
        lds_barrier();
        if (!valid)
            return;
#pragma unroll
        for (int k = 0; k < E / 2; ++k)
        {
            const int q    = ts + k * TP;
            const cx<T> uq = v[k];
            const cx<T> um = at(HN - 1 - q); // u[N - 1 - q]
            put<NT>(xq + q, quad<T>{ uq.re, um.im, uq.im, um.re });
        }
    }
}

template <typename T, int L>
void dct3_launch(const T* y, T* x, const cx<T>* tw, const cx<T>* wn, const cx<T>* w4n, long long count, hipStream_t s)
{
    using G                  = geom<T, L, role_direct>;
    static const bool lds_ok = (allow_lds(k_dct3<T, L>, G::lds_bytes), true);
    (void)lds_ok;
    const long long nb = (count + G::W - 1) / G::W;
    if (nb > 0x7fffffffll)
        throw std::runtime_error("batch too large for one launch");
    timing::span ts([] { return timing::kernel_name("k_dct3", timing::type_name<T>(), L); },
                    2 * (long long)sizeof(T) * (count << (L + 1)), s);
    hipLaunchKernelGGL((k_dct3<T, L>), dim3((unsigned)nb), dim3(G::BS), G::lds_bytes, s, y, x, tw, wn, w4n, count);
    hip_check(hipGetLastError(), "fused DCT-III launch");
}

template <typename T, int L>
void dct2_launch(const T* x, T* y, const cx<T>* tw, const cx<T>* wn, const cx<T>* w4n, long long count, hipStream_t s)
{
    using G                  = geom<T, L, role_direct>;
    static const bool lds_ok = (allow_lds(k_dct2<T, L>, G::lds_bytes), true);
    (void)lds_ok;
    const long long nb = (count + G::W - 1) / G::W;
    if (nb > 0x7fffffffll)
        throw std::runtime_error("batch too large for one launch");
    timing::span ts([] { return timing::kernel_name("k_dct2", timing::type_name<T>(), L); },
                    2 * (long long)sizeof(T) * (count << (L + 1)), s);
    hipLaunchKernelGGL((k_dct2<T, L>), dim3((unsigned)nb), dim3(G::BS), G::lds_bytes, s, x, y, tw, wn, w4n, count);
    hip_check(hipGetLastError(), "fused DCT launch");
}

// shortest length with a fused DCT: transforms of at least 16 threads (no staged short-transform loads)
template <typename T>
inline constexpr int dct_min_log = sizeof(T) == 4 ? 8 : 7;

template <typename T, int LO, int HI>
bool dct2_any(bool inverse, int L, const T* x, T* y, const cx<T>* tw, const cx<T>* wn, const cx<T>* w4n,
              long long count, hipStream_t s)
{
    if constexpr (LO <= HI)
    {
        if (L == LO)
        {
            if (inverse)
                dct3_launch<T, LO>(x, y, tw, wn, w4n, count, s);
            else
                dct2_launch<T, LO>(x, y, tw, wn, w4n, count, s);
            return true;
        }
        return dct2_any<T, LO + 1, HI>(inverse, L, x, y, tw, wn, w4n, count, s);
    }
    else
        return false;
}

// ---- Real transforms in one kernel (dft_plan_real, to_fmt / from_fmt of src/dft/fft-impl.hpp:1799-1896)
// for half lengths N = 2^L that fit one workgroup. The n = 2N reals are N complex values z.
// R2C: Z = DFT_N(z) in registers (slot k of thread t holds Z[t + k TPT]); one LDS round of N/2 values
// brings Z[N - i] to every slot i < N/2, which writes the outputs of both i and N - i
//   X[i] = (F1 + F2 w_i) / 2                     for 0 < i < N/2
//   X[i] = conj((F1' - F2' w_{N-i}) / 2)         for N/2 < i < N, primes for the pair (N - i, i)
// with F1 = Z[i] + conj Z[N-i], F2 = Z[i] - conj Z[N-i], w_k = rtw[k] = exp(-i(2 pi k/n + pi/2));
// X[0] and X[N] (CCs) or X[0] packed (Perm) from Z[0], X[N/2] = conj Z[N/2].
// C2R: each slot loads X[i]; the lower half of the slots receives X[N - i] through LDS, forms the
// packed inputs of both slots i and N - i, and returns the latter the same way; the inverse transform
// runs, and the reals are stored as plain pairs.
template <typename T, int L>
__global__ void __launch_bounds__((geom<T, L, role_direct>::BS), (geom<T, L, role_direct>::lds_waves))
k_rfft_fwd(const cx<T>* __restrict__ x, cx<T>* __restrict__ out, long long odist, const cx<T>* __restrict__ tw,
           const cx<T>* __restrict__ rtw, int fmt, long long count)
{
    using G          = geom<T, L, role_direct>;
    using S          = typename G::S;
    constexpr int N  = S::N;
    constexpr int E  = S::E;
    constexpr int TP = S::TPT;
    constexpr int W  = G::W;
    constexpr int HN = N / 2;
    static_assert(!G::staged && E % 2 == 0 && G::H <= 2, "fused real transform needs whole-thread transforms");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const mfma16 K{};
    const int tid     = threadIdx.x;
    const int tb      = tid / TP;
    const int t       = tid % TP;
    const long long g = (long long)blockIdx.x * W + tb;
    const bool valid  = g < count;
    auto at           = [&](int i) -> cx<T>& { return lds[laddr<role_direct, W, G::LD>(tb, i)]; };
    cx<T> v[E];
    const cx<T>* xb = x + g * N + t;
#pragma unroll
    for (int k = 0; k < E; ++k)
        v[k] = valid ? xb[k * TP] : cx<T>{ T(0), T(0) };
    run_pass<T, L, false, role_direct, prec_full, 0>(v, lds, tb, fresh(t), tw, K);
    cx<T>* o  = out + g * odist;
    // X[0] (and X[N]) from Z[0]
    auto emit_ends = [&](cx<T> a)
    {
        if (fmt == 1)
        {
            o[0] = { a.re + a.im, T(0) };
            o[N] = { a.re - a.im, T(0) };
        }
        else
            o[0] = { a.re + a.im, a.re - a.im };
    };
    // X[i] for 0 < i < N/2 (a = Z[i], c = Z[N-i])
    auto emit_lo = [&](int i, cx<T> a, cx<T> c)
    {
        const cx<T> f1 = { a.re + c.re, a.im - c.im }; // a + conj c
        const cx<T> f2 = { a.re - c.re, a.im + c.im }; // a - conj c
        const cx<T> tt = cmul(f2, rtw[i]);
        o[i]           = { T(0.5) * (f1.re + tt.re), T(0.5) * (f1.im + tt.im) };
    };
    // X[i] for N/2 < i < N: the pair (k, N - k) with k = N - i, Z[k] = c, Z[N - k] = a
    auto emit_hi = [&](int i, cx<T> a, cx<T> c)
    {
        const cx<T> f1 = { c.re + a.re, c.im - a.im };
        const cx<T> f2 = { c.re - a.re, c.im + a.im };
        const cx<T> tt = cmul(f2, rtw[N - i]);
        o[i]           = { T(0.5) * (f1.re - tt.re), -T(0.5) * (f1.im - tt.im) };
    };
    // the upper slots publish Z; each lower slot i reads Z[N-i] and writes the outputs of both i and
    // N - i (the second run in reverse order, still coalesced); slot N/2 pairs with itself. The
    // indices come from a fresh copy of t, so none is held through the passes.
    const int ts = fresh(t);
#pragma unroll
    for (int k = E / 2; k < E; ++k)
        at(ts + k * TP - HN) = v[k];
    lds_barrier();
    if (!valid)
        return;
#pragma unroll
    for (int k = 0; k < E / 2; ++k)
    {
        const int i = ts + k * TP;
        if (k == 0 && ts == 0)
            emit_ends(v[k]);
        else
        {
            const cx<T> c = at(N - i - HN);
            emit_lo(i, v[k], c);
            emit_hi(N - i, c, v[k]);
        }
    }
    if (ts == 0)
        o[HN] = cconj(v[E / 2]); // X[N/2] = conj Z[N/2]
}

template <typename T, int L>
__global__ void __launch_bounds__((geom<T, L, role_direct>::BS), (geom<T, L, role_direct>::lds_waves))
k_rfft_inv(const cx<T>* __restrict__ spec, long long idist, cx<T>* __restrict__ y, const cx<T>* __restrict__ tw,
           const cx<T>* __restrict__ rtw, int fmt, long long count)
{
    using G          = geom<T, L, role_direct>;
    using S          = typename G::S;
    constexpr int N  = S::N;
    constexpr int E  = S::E;
    constexpr int TP = S::TPT;
    constexpr int W  = G::W;
    static_assert(!G::staged && E % 2 == 0 && G::H <= 2, "fused real transform needs whole-thread transforms");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const mfma16 K{};
    const int tid     = threadIdx.x;
    const int tb      = tid / TP;
    const int t       = tid % TP;
    const long long g = (long long)blockIdx.x * W + tb;
    const bool valid  = g < count;
    const cx<T>* X   = spec + g * idist;
    constexpr int HN  = N / 2;
    auto at           = [&](int i) -> cx<T>& { return lds[laddr<role_direct, W, G::LD>(tb, i)]; };
    // the packed input of slot i from X[i] and X[N - i] (from_fmt)
    auto unpack = [&](int i, cx<T> xk, cx<T> xn) -> cx<T>
    {
        if (i == 0)
            return fmt == 1 ? cx<T>{ xk.re + xn.re, xk.re - xn.re } : cx<T>{ xk.re + xk.im, xk.re - xk.im };
        if (2 * i == N)
            return { T(2) * xk.re, T(-2) * xk.im };
        if (2 * i < N)
        {
            const cx<T> f1 = { xk.re + xn.re, xk.im - xn.im };
            const cx<T> f2 = { xk.re - xn.re, xk.im + xn.im };
            const cx<T> tt = cmulc(f2, rtw[i]);
            return { f1.re + tt.re, f1.im + tt.im };
        }
        // pair (k', N - k') with k' = N - i: X[k'] = xn, X[N - k'] = xk
        const cx<T> f1 = { xn.re + xk.re, xn.im - xk.im };
        const cx<T> f2 = { xn.re - xk.re, xn.im + xk.im };
        const cx<T> tt = cmulc(f2, rtw[N - i]);
        return { f1.re - tt.re, -(f1.im - tt.im) };
    };
    cx<T> v[E];
#pragma unroll
    for (int k = 0; k < E; ++k)
        v[k] = valid ? X[t + k * TP] : cx<T>{ T(0), T(0) };
    const cx<T> x0p = valid ? X[fmt == 1 ? N : 0] : cx<T>{ T(0), T(0) };
    // the upper slots publish X; each lower slot i reads X[N-i], forms both its own value and that of
    // slot N-i, and leaves the latter in the position it read (no other thread reads or writes it);
    // the upper slots collect theirs after one barrier. Slot 0 pairs with X[N] (CCs) or X[0] (Perm),
    // slot N/2 with itself.
#pragma unroll
    for (int k = E / 2; k < E; ++k)
        at(t + k * TP - HN) = v[k];
    lds_barrier();
#pragma unroll
    for (int k = 0; k < E / 2; ++k)
    {
        const int i = t + k * TP;
        if (i == 0)
            v[k] = unpack(0, v[k], x0p);
        else
        {
            cx<T>& slot    = at(N - i - HN);
            const cx<T> xp = slot;
            slot           = unpack(N - i, xp, v[k]);
            v[k]           = unpack(i, v[k], xp);
        }
    }
    lds_barrier();
#pragma unroll
    for (int k = E / 2; k < E; ++k)
    {
        const int i = t + k * TP;
        v[k]        = i == HN ? unpack(HN, v[k], v[k]) : at(i - HN);
    }
    lds_barrier();
    run_pass<T, L, true, role_direct, prec_full, 0>(v, lds, tb, t, tw, K);
    if (valid)
    {
        cx<T>* yb = y + g * N + t;
#pragma unroll
        for (int k = 0; k < E; ++k)
            yb[k * TP] = v[k];
    }
}

template <typename T, int L, bool INV>
void rfft_launch(const cx<T>* in, long long idist, cx<T>* out, long long odist, const cx<T>* tw, const cx<T>* rtw,
                 int fmt, long long count, hipStream_t s)
{
    using G = geom<T, L, role_direct>;
    if constexpr (INV)
    {
        static const bool lds_ok = (allow_lds(k_rfft_inv<T, L>, G::lds_bytes), true);
        (void)lds_ok;
    }
    else
    {
        static const bool lds_ok = (allow_lds(k_rfft_fwd<T, L>, G::lds_bytes), true);
        (void)lds_ok;
    }
    const long long nb = (count + G::W - 1) / G::W;
    if (nb > 0x7fffffffll)
        throw std::runtime_error("batch too large for one launch");
    // algorithmic bytes: the n = 2N reals one way, N + 1 (CCs) or N (Perm) complex values the other
    const long long bytes = (long long)sizeof(cx<T>) * count * ((2ll << L) + (fmt == 1 ? 1 : 0));
    timing::span ts([] { return timing::kernel_name(INV ? "k_rfft_inv" : "k_rfft_fwd", timing::type_name<T>(), L); },
                    bytes, s);
    if constexpr (INV)
        hipLaunchKernelGGL((k_rfft_inv<T, L>), dim3((unsigned)nb), dim3(G::BS), G::lds_bytes, s, in, idist, out, tw,
                           rtw, fmt, count);
    else
        hipLaunchKernelGGL((k_rfft_fwd<T, L>), dim3((unsigned)nb), dim3(G::BS), G::lds_bytes, s, in, out, odist, tw,
                           rtw, fmt, count);
    hip_check(hipGetLastError(), "fused real transform launch");
}

template <typename T, bool INV, int LO, int HI>
bool rfft_any(int L, const cx<T>* in, long long idist, cx<T>* out, long long odist, const cx<T>* tw,
              const cx<T>* rtw, int fmt, long long count, hipStream_t s)
{
    if constexpr (LO <= HI)
    {
        if (L == LO)
        {
            rfft_launch<T, LO, INV>(in, idist, out, odist, tw, rtw, fmt, count, s);
            return true;
        }
        return rfft_any<T, INV, LO + 1, HI>(L, in, idist, out, odist, tw, rtw, fmt, count, s);
    }
    else
        return false;
}

// ---- FFT convolution by overlap-save, one kernel per call.
// Output m of the linear convolution y = u * f (|f| = nf) needs u[m - nf + 1 .. m]. A segment of
// lv = N - nf + 1 consecutive outputs therefore reads N consecutive inputs z[n] = u[i0 + n], and the
// circular convolution of z with f zero-padded to N equals the linear one at n >= nf - 1. Each
// workgroup takes two adjacent segments of one signal as the real and imaginary parts of one complex
// sequence (f is real, so both convolutions come out separated), runs the forward transform, the
// product with the filter spectrum, and the inverse transform in registers and LDS, and writes
// the valid outputs: every signal value is read (N / lv) times and every output written once.
template <typename T>
struct conv_os_args
{
    const T* u       = nullptr; // input signals, item b at b * du
    T* y             = nullptr; // outputs, item b at b * dy
    const cx<T>* H   = nullptr; // filter spectra (N values each), item b's at b * hd (hd = 0: shared)
    const cx<T>* tw  = nullptr; // pass tables of N points
    long long du     = 0, dy = 0, hd = 0;
    long long nu     = 0; // input length
    long long o0     = 0; // full-convolution index of output 0
    long long nout   = 0; // outputs per item
    long long pairs  = 0; // segment pairs per item
    long long count  = 0; // items * pairs
    int nf           = 1; // filter length
    int lv           = 1; // valid outputs per segment, N - nf + 1
    T scale          = T(1); // 1 / N
};

template <typename T, int L>
__global__ void __launch_bounds__((geom<T, L, role_block>::BS)) k_conv_os(conv_os_args<T> a)
{
    using S          = typename geom<T, L, role_block>::S;
    constexpr int E  = S::E;
    constexpr int TP = S::TPT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const mfma16 K{};
    const long long w = blockIdx.x;
    if (w >= a.count)
        return;
    const long long item = w / a.pairs;
    const long long pair = w - item * a.pairs;
    const int t          = threadIdx.x;
    const T* u           = a.u + item * a.du;
    const long long sa   = 2 * pair;
    const long long ia   = a.o0 + sa * a.lv - (a.nf - 1); // input index of z[0], first segment
    const long long ib   = ia + a.lv;                      // second segment
    cx<T> v[E];
#pragma unroll
    for (int k = 0; k < E; ++k)
    {
        const long long n = t + k * TP;
        const long long pa = ia + n, pb = ib + n;
        v[k].re = (pa >= 0 && pa < a.nu) ? u[pa] : T(0);
        v[k].im = (pb >= 0 && pb < a.nu) ? u[pb] : T(0);
    }
    run_pass<T, L, false, role_block, prec_full, 0>(v, lds, 0, t, a.tw, K);
    // v[k] = Z[t + k*TPT]
    const cx<T>* H = a.H + item * a.hd;
#pragma unroll
    for (int k = 0; k < E; ++k)
    {
        const cx<T> h = H[t + k * TP];
        const cx<T> p = cmul(v[k], h);
        v[k]          = { p.re * a.scale, p.im * a.scale };
    }
    run_pass<T, L, true, role_block, prec_full, 0>(v, lds, 0, t, a.tw, K);
    T* y = a.y + item * a.dy;
#pragma unroll
    for (int k = 0; k < E; ++k)
    {
        const int n = t + k * TP;
        if (n >= a.nf - 1)
        {
            const long long j  = n - (a.nf - 1);
            const long long ma = sa * a.lv + j;
            const long long mb = ma + a.lv;
            if (ma < a.nout)
                y[ma] = v[k].re;
            if (mb < a.nout)
                y[mb] = v[k].im;
        }
    }
}

// Partial cross-correlations for the filter gradient: segment s of length P = N - nh + 1 of item b
// gives part[(b*nseg + s)*nh + j] = sum_{i<P} g[sP + i + j] x[sP + i], j < nh. One complex transform
// of z = x_s + i g_s (x zero beyond P) gives both spectra, X = (Z(k) + conj Z(-k))/2 and
// G = (Z(k) - conj Z(-k))/(2i); the inverse transform of G conj(X) is the circular correlation, which
// does not wrap for j < nh because P + nh - 1 = N.
template <typename T>
struct xcorr_args
{
    const T* x      = nullptr; // item b at b * dx, length nx
    const T* g      = nullptr; // item b at b * dg, length ng
    T* part         = nullptr;
    const cx<T>* tw = nullptr;
    long long dx = 0, dg = 0, nx = 0, ng = 0;
    long long nseg  = 0, count = 0;
    int P = 1, nh = 1;
    T scale = T(1); // 1 / N
};

template <typename T, int L>
__global__ void __launch_bounds__((geom<T, L, role_block>::BS)) k_conv_xcorr(xcorr_args<T> a)
{
    using G          = geom<T, L, role_block>;
    using S          = typename G::S;
    constexpr int N  = S::N;
    constexpr int E  = S::E;
    constexpr int TP = S::TPT;
    constexpr int HN = N / 2;
    static_assert(E % 2 == 0 && G::H <= 2, "mirror exchange in halves");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const mfma16 K{};
    const long long w = blockIdx.x;
    if (w >= a.count)
        return;
    const long long item = w / a.nseg;
    const long long seg  = w - item * a.nseg;
    const int t          = threadIdx.x;
    const T* x           = a.x + item * a.dx;
    const T* g           = a.g + item * a.dg;
    const long long base = seg * a.P;
    auto at              = [&](int i) -> cx<T>& { return lds[laddr<role_block, 1, G::LD>(0, i)]; };
    cx<T> v[E];
#pragma unroll
    for (int k = 0; k < E; ++k)
    {
        const int n       = t + k * TP;
        const long long p = base + n;
        v[k].re           = (n < a.P && p < a.nx) ? x[p] : T(0);
        v[k].im           = p < a.ng ? g[p] : T(0);
    }
    run_pass<T, L, false, role_block, prec_full, 0>(v, lds, 0, t, a.tw, K);
    // v[k] = Z[i], i = t + k*TPT. R[i] = G[i] conj(X[i]) / N needs Z[N - i]. X and G are spectra of
    // real signals, so R[N - i] = conj(R[i]): the upper slots (i >= N/2) publish Z, each lower slot
    // reads its partner, forms R[i] and leaves conj(R[i]) = R[N - i] in the position it read, and the
    // upper slots collect theirs (slot 0 pairs with itself, as does N/2).
    auto prod = [&](cx<T> z, cx<T> zm) -> cx<T>
    {
        const cx<T> X   = { T(0.5) * (z.re + zm.re), T(0.5) * (z.im - zm.im) };
        const cx<T> dgi = { z.re - zm.re, z.im + zm.im }; // G = (z - conj zm) / (2i)
        const cx<T> Gs  = { T(0.5) * dgi.im, -T(0.5) * dgi.re };
        return { (Gs.re * X.re + Gs.im * X.im) * a.scale, (Gs.im * X.re - Gs.re * X.im) * a.scale };
    };
#pragma unroll
    for (int k = E / 2; k < E; ++k)
        at(t + k * TP - HN) = v[k];
    lds_barrier();
#pragma unroll
    for (int k = 0; k < E / 2; ++k)
    {
        const int i = t + k * TP;
        if (i == 0)
            v[k] = prod(v[k], v[k]);
        else
        {
            cx<T>& slot = at(N - i - HN);
            v[k]        = prod(v[k], slot);
            slot        = cconj(v[k]); // R[N - i], for the slot that published Z[N - i]
        }
    }
    lds_barrier();
#pragma unroll
    for (int k = E / 2; k < E; ++k)
    {
        const int i = t + k * TP;
        v[k]        = i == HN ? prod(v[k], v[k]) : at(i - HN);
    }
    lds_barrier();
    run_pass<T, L, true, role_block, prec_full, 0>(v, lds, 0, t, a.tw, K);
    T* out = a.part + w * a.nh;
#pragma unroll
    for (int k = 0; k < E; ++k)
    {
        const int n = t + k * TP;
        if (n < a.nh)
            out[n] = v[k].re;
    }
}

// out[b][j] = sum over s < nseg of part[b][s][j] (a shared filter sums the items in a second call)
template <typename T>
__global__ void k_conv_xcorr_sum(T* out, const T* part, long long nh, long long nseg, long long items)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= items * nh)
        return;
    const long long b = i / nh;
    const long long j = i - b * nh;
    const T* p        = part + b * nseg * nh + j;
    T acc             = T(0);
    for (long long s = 0; s < nseg; ++s)
        acc += p[s * nh];
    out[i] = acc;
}

// complex filter buffer (f[n], 0) for n < nf, zero to N; f = h reversed when rev
template <typename T>
__global__ void k_conv_filter_pad(cx<T>* out, const T* h, long long hd, int nf, int N, long long count, bool rev)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count * N)
        return;
    const long long b = i / N;
    const int n       = (int)(i - b * N);
    const T* hb       = h + b * hd;
    out[i]            = { n < nf ? hb[rev ? nf - 1 - n : n] : T(0), T(0) };
}

template <typename T, int L>
void conv_os_launch(const conv_os_args<T>& a, hipStream_t s)
{
    using G                  = geom<T, L, role_block>;
    constexpr size_t bytes   = G::lds_bytes;
    static const bool lds_ok = (allow_lds(k_conv_os<T, L>, bytes), true);
    (void)lds_ok;
    if (a.count > 0x7fffffffll)
        throw std::runtime_error("convolution too large for one launch");
    timing::span ts([] { return timing::kernel_name("k_conv_os", timing::type_name<T>(), L); },
                    (long long)sizeof(T) * (a.count / a.pairs) * (a.nu + a.nout), s);
    hipLaunchKernelGGL((k_conv_os<T, L>), dim3((unsigned)a.count), dim3(G::BS), bytes, s, a);
    hip_check(hipGetLastError(), "overlap-save convolution launch");
}

template <typename T, int L>
void conv_xcorr_launch(const xcorr_args<T>& a, hipStream_t s)
{
    using G                  = geom<T, L, role_block>;
    constexpr size_t bytes   = G::lds_bytes;
    static const bool lds_ok = (allow_lds(k_conv_xcorr<T, L>, bytes), true);
    (void)lds_ok;
    if (a.count > 0x7fffffffll)
        throw std::runtime_error("correlation too large for one launch");
    timing::span ts([] { return timing::kernel_name("k_conv_xcorr", timing::type_name<T>(), L); },
                    (long long)sizeof(T) * (a.count / a.nseg) * (a.nx + a.ng) +
                        (long long)sizeof(T) * a.count * a.nh,
                    s);
    hipLaunchKernelGGL((k_conv_xcorr<T, L>), dim3((unsigned)a.count), dim3(G::BS), bytes, s, a);
    hip_check(hipGetLastError(), "correlation launch");
}

constexpr int conv_min_log = 8;

template <typename T, int LO>
void conv_os_any(int L, const conv_os_args<T>& a, hipStream_t s)
{
    if constexpr (LO <= max_log<T>)
    {
        if (L == LO)
            return conv_os_launch<T, LO>(a, s);
        return conv_os_any<T, LO + 1>(L, a, s);
    }
    else
        throw std::logic_error("no overlap-save kernel for this block length");
}

template <typename T, int LO>
void conv_xcorr_any(int L, const xcorr_args<T>& a, hipStream_t s)
{
    if constexpr (LO <= max_log<T>)
    {
        if (L == LO)
            return conv_xcorr_launch<T, LO>(a, s);
        return conv_xcorr_any<T, LO + 1>(L, a, s);
    }
    else
        throw std::logic_error("no correlation kernel for this block length");
}

} // namespace p2

template <typename T>
struct pow2_fft<T>::impl
{
    int L     = 0;
    int prec  = p2::prec_full;
    bool four = false;
    int L1 = 0, L2 = 0; // four-step: rows of n1 = 2^L1 points, columns of n2 = 2^L2 points
    int lbits = 0;
    dev_mem tw;
    long long o_row = 0, o_ta = 0, o_tb = 0;
    dev_mem scratch, scales;
    int gshift      = 0;
    long long chunk = 1;
    // multi-pass Stockham over HBM (long lengths): log2 radix of every pass, its sub-DFT tables and
    // its W_{Ns*R} two-level tables
    bool multi = false;
    std::vector<int> mp_log, mp_lbits;
    std::vector<long long> mp_tw, mp_ta, mp_tb;
    dev_mem scratch2;

    template <int PREC>
    void run(bool inverse, const void* in, void* out, long long batch, hipStream_t s, workspace w);
};

template <typename T>
bool pow2_fft<T>::supports(size_t n)
{
    const int L = p2::log2_exact(n);
    return L >= 1 && L <= 2 * p2::max_log<T>;
}

template <typename T>
pow2_fft<T>::pow2_fft(size_t n, hipStream_t s, int prec) : p(new impl)
{
    const int L = p2::log2_exact(n);
    if (!supports(n))
        throw std::invalid_argument("pow2_fft: unsupported length");
    if (prec != p2::prec_full && (sizeof(T) != 4 || (prec != p2::prec_bf16 && prec != p2::prec_f16)))
        throw std::invalid_argument("pow2_fft: half-precision storage is computed in f32");
    const int le = p2::log_e<T>;
    p->L         = L;
    p->prec      = prec;
    if (L <= (prec == p2::prec_full ? p2::direct_max_log<T> : p2::max_log<T>))
    {
        const int cnt = p2::pass_twoff(L, le, p2::pass_count(L, le));
        p->tw.reset((size_t)std::max(1, cnt) * sizeof(cx<T>));
        p2::gen_pass_tables<T>(L, p->tw.template as<cx<T>>(), s);
        return;
    }
    // the longest lengths (measured with 1 GiB rounds: from 2^24; at 2^23 the four-step is 10% (f32)
    // and 2% (f64) faster): HBM passes of radix 2^5..2^8 with 16 butterflies per workgroup (128-byte
    // runs in and out, small workgroups); shorter lengths: the two-kernel four-step, whose extra pass
    // costs more than its wider workgroups
    int mp_min = 24;
    if (const char* e = std::getenv("KFR_AMD_MULTIPASS_MIN_LOG"))
        mp_min = atoi(e);
    // half-precision storage: the multi-pass kernels (32 butterflies per workgroup, 128-byte runs of
    // 4-byte values) take 2^24 points, where they beat the four-step (0.94 against 1.00 ms per
    // 256 MiB of bf16 data); from 2^17 to 2^23 the four-step's two passes are 30-50% faster than
    // three multi-pass passes, whose kernels are bound by instructions, not bytes.
    // KFR_AMD_HALF_MULTIPASS_MIN_LOG moves the threshold.
    int half_mp_min = 24;
    if (const char* e = std::getenv("KFR_AMD_HALF_MULTIPASS_MIN_LOG"))
        half_mp_min = atoi(e);
    if ((prec == p2::prec_full && L >= mp_min) || (prec != p2::prec_full && L >= half_mp_min))
    {
        p->multi     = true;
        const int np = (L + 7) / 8;
        // shorter passes first: the last pass (largest stride, widest twiddle gathers) is the slowest,
        // so it gets a full radix-2^8 workgroup (measured 3% faster for f64 2^23)
        for (int q = 0; q < np; ++q)
            p->mp_log.push_back(L / np + (q >= np - L % np ? 1 : 0));
        long long off = 0;
        int nsl       = 0;
        for (int q = 0; q < np; ++q)
        {
            const int l = p->mp_log[q];
            p->mp_tw.push_back(off);
            off += (p2::pass_twoff(l, le, p2::pass_count(l, le)) + 1) / 2 * 2;
            const int s  = nsl + l; // the pass's twiddles are powers of W_{2^s}
            const int lb = (s + 1) / 2;
            p->mp_lbits.push_back(lb);
            p->mp_ta.push_back(off);
            off += ((1ll << s >> lb) + 1 + 1) / 2 * 2;
            p->mp_tb.push_back(off);
            off += ((1ll << lb) + 1) / 2 * 2;
            nsl += l;
        }
        p->tw.reset((size_t)off * sizeof(cx<T>));
        cx<T>* t = p->tw.template as<cx<T>>();
        nsl      = 0;
        for (int q = 0; q < np; ++q)
        {
            const int l       = p->mp_log[q];
            p2::gen_pass_tables<T>(l, t + p->mp_tw[q], s);
            const long long M = 1ll << (nsl + l);
            const int lb      = p->mp_lbits[q];
            gen_twiddles(t + p->mp_ta[q], (M >> lb) + 1, M, 1ll << lb, 0, tw_linear, 0, s);
            gen_twiddles(t + p->mp_tb[q], 1ll << lb, M, 1, 0, tw_linear, 0, s);
            nsl += l;
        }
        size_t cap = (size_t)1 << 30; // rounds of at most 1 GiB, as the four-step below
        if (const char* e = std::getenv("KFR_AMD_FOURSTEP_CHUNK_MB"))
            cap = (size_t)std::max(1, atoi(e)) << 20;
        const size_t mid = prec == p2::prec_full ? sizeof(cx<T>) : 2 * sizeof(_Float16);
        p->chunk         = (long long)std::max<size_t>(1, cap / (n * mid));
        return;
    }
    p->four = true;
    p->L2   = std::max(p2::min_split_log, L / 2);
    // column length 2^L2 measured fastest per length (sweep of L2 over 7..14 on MI355X, full
    // precision): the column and row kernels of 2^8..2^9 points run near half of peak, those of
    // 2^10..2^12 points near a third, so uneven splits often win
    if (prec == p2::prec_full)
    {
        static constexpr int best_f32[] = { 8, 8, 9, 9, 8, 8, 9, 11, 11 };  // L = 15..23
        static constexpr int best_f64[] = { 7, 9, 7, 8, 9, 8, 9, 11, 11 };  // L = 14..22
        if (sizeof(T) == 4 && L >= 15 && L <= 23)
            p->L2 = best_f32[L - 15];
        if (sizeof(T) == 8 && L >= 14 && L <= 22)
            p->L2 = best_f64[L - 14];
    }
    else if (L == 20)
        p->L2 = 12; // half precision, 2^20: measured 6% faster than 2^10 x 2^10
    if (const char* e = std::getenv("KFR_AMD_FOURSTEP_L2"))
        p->L2 = std::atoi(e);
    p->L2 = std::min(std::max(p->L2, std::max(p2::min_split_log, L - p2::max_log<T>)),
                     std::min(p2::max_log<T>, L - p2::min_split_log));
    p->L1 = L - p->L2;
    const int c2         = p2::pass_twoff(p->L2, le, p2::pass_count(p->L2, le));
    const int c1         = p2::pass_twoff(p->L1, le, p2::pass_count(p->L1, le));
    p->lbits             = (L + 1) / 2;
    const long long nTB  = 1ll << p->lbits;
    const long long nTA  = (long long)(n >> p->lbits) + 1;
    p->o_row             = (c2 + 1) / 2 * 2; // keep every table 16-byte aligned
    p->o_ta              = p->o_row + (c1 + 1) / 2 * 2;
    p->o_tb              = p->o_ta + (nTA + 1) / 2 * 2;
    p->tw.reset((size_t)(p->o_tb + nTB) * sizeof(cx<T>));
    cx<T>* t = p->tw.template as<cx<T>>();
    p2::gen_pass_tables<T>(p->L2, t, s);
    p2::gen_pass_tables<T>(p->L1, t + p->o_row, s);
    gen_twiddles(t + p->o_ta, nTA, (long long)n, nTB, 0, tw_linear, 0, s);
    gen_twiddles(t + p->o_tb, nTB, (long long)n, 1, 0, tw_linear, 0, s);
    // binary16 and bfloat16 storage share one geometry
    int w = prec == p2::prec_full ? p2::col_width<T, p2::prec_full, p2::min_split_log, p2::max_log<T>>(p->L2)
                                  : p2::col_width<T, p2::prec_bf16, p2::min_split_log, p2::max_log<T>>(p->L2);
    while ((1 << p->gshift) < w)
        ++p->gshift;
    // rounds of at most 1 GiB of intermediate: few launches and round boundaries (measured 4-16% faster
    // than 512 MiB rounds from 2^20 up, which were faster than rounds small enough to stay in the
    // last-level cache), bounded scratch
    size_t cap = (size_t)1 << 30;
    if (const char* e = std::getenv("KFR_AMD_FOURSTEP_CHUNK_MB"))
        cap = (size_t)std::max(1, atoi(e)) << 20;
    const size_t mid = prec == p2::prec_full ? sizeof(cx<T>) : 2 * sizeof(_Float16);
    p->chunk         = (long long)std::max<size_t>(1, cap / (n * mid));
}

template <typename T>
pow2_fft<T>::~pow2_fft() = default;

template <typename T>
template <int PREC>
void pow2_fft<T>::impl::run(bool inverse, const void* in, void* out, long long batch, hipStream_t s, workspace w)
{
    using p2::args;
    constexpr int ML = p2::max_log<T>;
    const size_t es  = PREC == p2::prec_full ? sizeof(cx<T>) : 2 * sizeof(_Float16); // HBM element size
    const cx<T>* t   = tw.template as<cx<T>>();
    if constexpr (PREC != p2::prec_full)
    {
        if (multi)
        {
            // binary16 intermediates, one power-of-two scale per workgroup, ping-pong between two
            // buffers (and two scale arrays)
            if (w.p)
                throw std::logic_error("half-precision transforms keep their own intermediate");
            const long long n = 1ll << L;
            const int np      = (int)mp_log.size();
            const int nbuf    = np > 2 ? 2 : 1;
            const long long c = std::min(chunk, batch);
            scratch.ensure((size_t)c * n * es);
            if (nbuf > 1)
                scratch2.ensure((size_t)c * n * es);
            long long maxwg = 1;
            for (int q = 0; q + 1 < np; ++q)
                maxwg = std::max(maxwg, ((c << (L - mp_log[q])) + 31) / p2::pass_width<T, PREC, 5, 8>(mp_log[q]) + 1);
            scales.ensure((size_t)(2 * maxwg) * sizeof(int));
            void* buf[2] = { scratch.ptr, scratch2.ptr };
            int* scb[2]  = { scales.template as<int>(), scales.template as<int>() + maxwg };
            for (long long b0 = 0; b0 < batch; b0 += chunk)
            {
                const long long nb = std::min(chunk, batch - b0);
                int nsl = 0, prev_l = 0, prev_ns = 0, prev_w = 0;
                for (int q = 0; q < np; ++q)
                {
                    const int l = mp_log[q];
                    args<T> a;
                    a.in        = q == 0 ? static_cast<const char*>(in) + b0 * n * es : buf[(q - 1) & 1];
                    a.out       = q == np - 1 ? static_cast<char*>(out) + b0 * n * es : buf[q & 1];
                    a.tw        = t + mp_tw[q];
                    a.ta        = t + mp_ta[q];
                    a.tb        = t + mp_tb[q];
                    a.lbits     = mp_lbits[q];
                    a.lo        = L - l;
                    a.ns_log    = nsl;
                    a.out_mid   = q + 1 < np ? 1 : 0;
                    a.scales    = q + 1 < np ? scb[q & 1] : nullptr;
                    a.in_scales = q > 0 ? scb[(q - 1) & 1] : nullptr;
                    a.in_l      = prev_l;
                    a.in_ns     = prev_ns;
                    a.in_wlog   = prev_w;
                    p2::launch_any<T, role_pass, PREC, 5, 8>(inverse, l, a, nb << (L - l), s);
                    prev_l  = l;
                    prev_ns = nsl;
                    prev_w  = p2::ilog2c(p2::pass_width<T, PREC, 5, 8>(l));
                    nsl += l;
                }
            }
            return;
        }
    }
    if constexpr (PREC == p2::prec_full)
    {
        if (multi)
        {
            const long long n     = 1ll << L;
            const int nbuf        = mp_log.size() > 2 ? 2 : 1;
            const long long chunk = rounds_in(w, this->chunk, (size_t)n * sizeof(cx<T>), nbuf);
            void* buf[2]          = { nullptr, nullptr };
            if (w.p)
            {
                buf[0] = workspace_part(w, 0, nbuf);
                buf[1] = nbuf > 1 ? workspace_part(w, 1, nbuf) : nullptr;
            }
            else
            {
                const size_t bytes = (size_t)std::min(chunk, batch) * n * sizeof(cx<T>);
                scratch.ensure(bytes);
                if (nbuf > 1)
                    scratch2.ensure(bytes);
                buf[0] = scratch.ptr;
                buf[1] = scratch2.ptr;
            }
            const int np = (int)mp_log.size();
            for (long long b0 = 0; b0 < batch; b0 += chunk)
            {
                const long long nb = std::min(chunk, batch - b0);
                int nsl            = 0;
                for (int q = 0; q < np; ++q)
                {
                    const int l = mp_log[q];
                    args<T> a;
                    a.in     = q == 0 ? static_cast<const char*>(in) + b0 * n * es : buf[(q - 1) & 1];
                    a.out    = q == np - 1 ? static_cast<char*>(out) + b0 * n * es : buf[q & 1];
                    a.tw     = t + mp_tw[q];
                    a.ta     = t + mp_ta[q];
                    a.tb     = t + mp_tb[q];
                    a.lbits  = mp_lbits[q];
                    a.lo     = L - l;
                    a.ns_log = nsl;
                    p2::launch_any<T, role_pass, PREC, 5, 8>(inverse, l, a, nb << (L - l), s);
                    nsl += l;
                }
            }
            return;
        }
    }
    if (!four)
    {
        args<T> a;
        a.in  = in;
        a.out = out;
        a.tw  = t;
        constexpr int DL = PREC == p2::prec_full ? p2::direct_max_log<T> : ML;
        p2::launch_any<T, role_direct, PREC, 1, DL>(inverse, L, a, batch, s);
        return;
    }
    const long long n = 1ll << L;
    const size_t mid  = PREC == p2::prec_full ? sizeof(cx<T>) : 2 * sizeof(_Float16);
    if (w.p && PREC != p2::prec_full)
        throw std::logic_error("half-precision transforms keep their own intermediate");
    const long long chunk = rounds_in(w, this->chunk, (size_t)n * mid, 1);
    void* midbuf          = w.p;
    if (!w.p)
    {
        scratch.ensure((size_t)std::min(chunk, batch) * n * mid);
        midbuf = scratch.ptr;
    }
    if (PREC != p2::prec_full)
        scales.ensure((size_t)(std::min(chunk, batch) << L1 >> gshift) * sizeof(int));
    for (long long b0 = 0; b0 < batch; b0 += chunk)
    {
        const long long nb = std::min(chunk, batch - b0);
        args<T> a;
        a.ta     = t + o_ta;
        a.tb     = t + o_tb;
        a.lbits  = lbits;
        a.scales = scales.template as<int>();
        a.gshift = gshift;
        a.in     = static_cast<const char*>(in) + b0 * n * es;
        a.out    = midbuf;
        a.tw     = t;
        a.lo     = L1;
        p2::launch_any<T, role_col, PREC, p2::min_split_log, ML>(inverse, L2, a, nb << L1, s);
        a.in  = midbuf;
        a.out = static_cast<char*>(out) + b0 * n * es;
        a.tw  = t + o_row;
        a.lo  = L2;
        p2::launch_any<T, role_row, PREC, p2::min_split_log, ML>(inverse, L1, a, nb << L2, s);
    }
}

template <typename T>
void pow2_fft<T>::run(bool inverse, const void* in, void* out, long long batch, hipStream_t s, workspace w)
{
    if (batch <= 0)
        return;
    if constexpr (sizeof(T) == 4)
    {
        if (p->prec == p2::prec_bf16)
            return p->template run<p2::prec_bf16>(inverse, in, out, batch, s, w);
        if (p->prec == p2::prec_f16)
            return p->template run<p2::prec_f16>(inverse, in, out, batch, s, w);
    }
    p->template run<p2::prec_full>(inverse, in, out, batch, s, w);
}

template <typename T>
bool pow2_fft<T>::supports_axis(size_t n)
{
    const int L = p2::log2_exact(n);
    return L >= p2::min_split_log && L <= p2::max_log<T>;
}

template <typename T>
void pow2_fft<T>::run_axis(bool inverse, const void* in, void* out, int inner_log, long long outer, hipStream_t s)
{
    if (!supports_axis(1ull << p->L) || p->prec != p2::prec_full || p->four || p->multi)
        throw std::logic_error("pow2_fft: this plan cannot transform a strided axis");
    if (outer <= 0)
        return;
    p2::args<T> a;
    a.in  = in;
    a.out = out;
    a.tw  = p->tw.template as<cx<T>>();
    a.ta  = nullptr; // no four-step twiddles
    a.lo  = inner_log;
    constexpr int ML = p2::max_log<T>;
    p2::launch_any<T, role_col, p2::prec_full, p2::min_split_log, ML>(inverse, p->L, a, outer << inner_log, s);
}

template <typename T>
bool pow2_fft<T>::dct(bool inverse, const void* x, void* y, const void* wn, const void* w4n, long long batch,
                      hipStream_t s)
{
    if (p->four || p->multi || p->prec != p2::prec_full || p->L < p2::dct_min_log<T> ||
        p->L > p2::direct_max_log<T>)
        return false;
    if (batch <= 0)
        return true;
    return p2::dct2_any<T, p2::dct_min_log<T>, p2::direct_max_log<T>>(
        inverse, p->L, static_cast<const T*>(x), static_cast<T*>(y), p->tw.template as<cx<T>>(),
        static_cast<const cx<T>*>(wn), static_cast<const cx<T>*>(w4n), batch, s);
}

template <typename T>
bool pow2_fft<T>::real(bool inverse, const void* in, long long idist, void* out, long long odist, const void* rtw,
                       int fmt, long long batch, hipStream_t s)
{
    if (p->four || p->multi || p->prec != p2::prec_full || p->L < p2::dct_min_log<T> ||
        p->L > p2::direct_max_log<T>)
        return false;
    if (batch <= 0)
        return true;
    const cx<T>* t = p->tw.template as<cx<T>>();
    const auto* i  = static_cast<const cx<T>*>(in);
    auto* o        = static_cast<cx<T>*>(out);
    const auto* r  = static_cast<const cx<T>*>(rtw);
    if (inverse)
        return p2::rfft_any<T, true, p2::dct_min_log<T>, p2::direct_max_log<T>>(p->L, i, idist, o, odist, t, r, fmt,
                                                                                 batch, s);
    return p2::rfft_any<T, false, p2::dct_min_log<T>, p2::direct_max_log<T>>(p->L, i, idist, o, odist, t, r, fmt, batch,
                                                                              s);
}

template <typename T>
std::string pow2_fft<T>::describe() const
{
    const int le = p2::log_e<T>;
    auto radices = [&](int L)
    {
        std::string r;
        for (int q = 0; q < p2::pass_count(L, le); ++q)
            r += (q ? "x" : "") + std::to_string(1 << p2::pass_lr(L, le, q));
        return r;
    };
    const std::string pr = p->prec == p2::prec_bf16 ? " storage=bf16 radix16=mfma"
                           : p->prec == p2::prec_f16 ? " storage=f16 radix16=mfma"
                                                      : "";
    if (p->multi)
    {
        std::string r;
        for (size_t q = 0; q < p->mp_log.size(); ++q)
            r += (q ? "x" : "") + std::to_string(1 << p->mp_log[q]);
        const int w = p->prec == p2::prec_full ? 16 : 32;
        return "pow2 multi-pass radices=" + r + " (" + std::to_string(w) +
               " butterflies per workgroup, one HBM pass each) chunk=" + std::to_string(p->chunk) + pr +
               (p->prec != p2::prec_full ? " intermediate=f16 scaled per workgroup" : "");
    }
    if (!p->four)
        return "pow2 radices=" + radices(p->L) + " threads/transform=" +
               std::to_string(p2::direct_threads<T, 1, p2::direct_max_log<T>>(p->L)) + pr;
    return "pow2 four-step n1=" + std::to_string(1ll << p->L1) + " n2=" + std::to_string(1ll << p->L2) +
           " radices=" + radices(p->L2) + "|" + radices(p->L1) + " chunk=" + std::to_string(p->chunk) + pr +
           (p->prec != p2::prec_full ? " intermediate=f16 scaled per " + std::to_string(1 << p->gshift) + " columns"
                                     : "");
}

template class pow2_fft<float>;
template class pow2_fft<double>;

// ------------------------------------------------------------------------------------------------

template <typename T>
struct os_conv<T>::impl
{
    size_t nx = 0, nh = 0;
    int L     = 0;
    dev_mem tw, hz, H, part, sum;
    std::unique_ptr<pow2_fft<T>> fft; // filter spectra
};

template <typename T>
int os_conv<T>::block_log(size_t nx, size_t nh)
{
    (void)nx;
    if (nh == 0)
        return -1;
    // the shortest block whose valid part is at least two thirds of it (half at the longest block)
    for (int L = p2::conv_min_log; L <= p2::max_log<T>; ++L)
        if (3 * (nh - 1) <= ((size_t)1 << L))
            return L;
    return 2 * (nh - 1) <= ((size_t)1 << p2::max_log<T>) ? p2::max_log<T> : -1;
}

template <typename T>
os_conv<T>::os_conv(size_t nx, size_t nh, hipStream_t s) : p(new impl)
{
    const int L = block_log(nx, nh);
    if (L < 0)
        throw std::invalid_argument("os_conv: filter too long for an LDS block");
    p->nx = nx;
    p->nh = nh;
    p->L  = L;
    const int le  = p2::log_e<T>;
    const int cnt = p2::pass_twoff(L, le, p2::pass_count(L, le));
    p->tw.reset((size_t)std::max(1, cnt) * sizeof(cx<T>));
    p2::gen_pass_tables<T>(L, p->tw.template as<cx<T>>(), s);
    p->fft.reset(new pow2_fft<T>((size_t)1 << L, s));
}

template <typename T>
os_conv<T>::~os_conv() = default;

template <typename T>
std::string os_conv<T>::describe() const
{
    return "overlap-save block=" + std::to_string(1 << p->L) + " valid=" +
           std::to_string((1 << p->L) - (long long)p->nh + 1);
}

// spectra of the filters (reversed when rev) scaled later by 1/N inside the kernels
template <typename T>
static const cx<T>* filter_spectra(typename os_conv<T>::impl& q, const T* h, long long h_dist, long long nfilt,
                                   bool rev, hipStream_t s)
{
    const long long N = 1ll << q.L;
    q.hz.ensure((size_t)(nfilt * N) * sizeof(cx<T>));
    q.H.ensure((size_t)(nfilt * N) * sizeof(cx<T>));
    hipLaunchKernelGGL(p2::k_conv_filter_pad<T>, dim3(nblocks(nfilt * N, 256)), dim3(256), 0, s,
                       q.hz.template as<cx<T>>(), h, h_dist, (int)q.nh, (int)N, nfilt, rev);
    hip_check(hipGetLastError(), "filter padding");
    q.fft->run(false, q.hz.ptr, q.H.ptr, nfilt, s);
    return q.H.template as<cx<T>>();
}

template <typename T>
static void run_os(typename os_conv<T>::impl& q, T* y, const T* u, long long nu, long long du, long long o0,
                   long long nout, long long dy, const cx<T>* H, long long hd, long long batch, hipStream_t s)
{
    const int N = 1 << q.L;
    p2::conv_os_args<T> a;
    a.u     = u;
    a.y     = y;
    a.H     = H;
    a.tw    = q.tw.template as<cx<T>>();
    a.du    = du;
    a.dy    = dy;
    a.hd    = hd;
    a.nu    = nu;
    a.o0    = o0;
    a.nout  = nout;
    a.nf    = (int)q.nh;
    a.lv    = N - (int)q.nh + 1;
    const long long nseg = (nout + a.lv - 1) / a.lv;
    a.pairs = (nseg + 1) / 2;
    a.count = a.pairs * batch;
    a.scale = T(1) / T(N);
    if (a.count > 0)
        p2::conv_os_any<T, p2::conv_min_log>(q.L, a, s);
}

template <typename T>
void os_conv<T>::forward(T* y, const T* x, const T* h, long long h_dist, long long batch, bool correlate,
                         hipStream_t s)
{
    if (batch <= 0)
        return;
    impl& q              = *p;
    const long long nfil = h_dist == 0 ? 1 : batch;
    const cx<T>* H       = filter_spectra<T>(q, h, h_dist, nfil, correlate, s);
    const long long ny   = (long long)(q.nx + q.nh - 1);
    run_os<T>(q, y, x, (long long)q.nx, (long long)q.nx, 0, ny, ny, H, h_dist == 0 ? 0 : (1ll << q.L), batch, s);
}

template <typename T>
void os_conv<T>::backward(T* dx, T* dh, const T* g, const T* x, const T* h, long long h_dist, long long batch,
                          bool correlate, hipStream_t s)
{
    if (batch <= 0)
        return;
    impl& q            = *p;
    const long long ny = (long long)(q.nx + q.nh - 1);
    const long long N  = 1ll << q.L;
    if (dx)
    {
        // dx[n] = sum_j g[n + j] f[j] = (g * reversed f)[n + nh - 1], f = h (convolution) or h reversed
        // (correlation)
        const long long nfil = h_dist == 0 ? 1 : batch;
        const cx<T>* H       = filter_spectra<T>(q, h, h_dist, nfil, !correlate, s);
        run_os<T>(q, dx, g, ny, ny, (long long)q.nh - 1, (long long)q.nx, (long long)q.nx, H,
                  h_dist == 0 ? 0 : N, batch, s);
    }
    if (dh)
    {
        p2::xcorr_args<T> a;
        a.x     = x;
        a.g     = g;
        a.tw    = q.tw.template as<cx<T>>();
        a.dx    = (long long)q.nx;
        a.dg    = ny;
        a.nx    = (long long)q.nx;
        a.ng    = ny;
        a.P     = (int)(N - (long long)q.nh + 1);
        a.nh    = (int)q.nh;
        a.nseg  = ((long long)q.nx + a.P - 1) / a.P;
        a.count = a.nseg * batch;
        a.scale = T(1) / T(N);
        q.part.ensure((size_t)(a.count * a.nh) * sizeof(T));
        a.part = q.part.template as<T>();
        if (a.count > 0)
            p2::conv_xcorr_any<T, p2::conv_min_log>(q.L, a, s);
        // per item over its segments, then (shared filter) over the items
        const long long nh = (long long)q.nh;
        T* per_item        = dh;
        if (h_dist == 0)
        {
            q.sum.ensure((size_t)(batch * nh) * sizeof(T));
            per_item = q.sum.template as<T>();
        }
        hipLaunchKernelGGL(p2::k_conv_xcorr_sum<T>, dim3(nblocks(batch * nh, 256)), dim3(256), 0, s, per_item,
                           a.part, nh, a.nseg, batch);
        hip_check(hipGetLastError(), "gradient reduction");
        if (h_dist == 0)
        {
            hipLaunchKernelGGL(p2::k_conv_xcorr_sum<T>, dim3(nblocks(nh, 256)), dim3(256), 0, s, dh, per_item, nh,
                               batch, 1ll);
            hip_check(hipGetLastError(), "gradient reduction");
        }
    }
}

template class os_conv<float>;
template class os_conv<double>;

} // namespace kfr_amd
