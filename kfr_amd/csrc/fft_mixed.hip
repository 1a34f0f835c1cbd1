// Mixed-radix C2C kernels specialised per length (f32).
//
// A transform of N points (a menu length) runs its Stockham passes with compile-time radices (16,
// then the power-of-two remainder, then 7, 5, 3), strides and loop bounds. Each pass gives every thread
// ceil(16/R) butterflies of radix R; butterfly j reads x[j + r*N/R]. The first pass reads those
// operands straight from HBM (consecutive threads, consecutive addresses), passes exchange through
// padded LDS, and the last pass, whose output r of butterfly j is element j + r*N/R in natural order,
// writes straight to HBM. Longer lengths use the four-step split n = n1*n2 with both factors from the
// menu: the column kernel transforms W adjacent columns per workgroup and applies W_n^(c*k2), the row
// kernel stores its rows transposed through LDS; the other dimension is a run-time divisor (fastdiv).
#include "fft_mixed.hpp"
#include "fft_timing.hpp"

#include "fft_core.hpp"

#include <array>
#include <stdexcept>
#include <utility>

namespace kfr_amd
{
namespace mx
{

constexpr int menu[] = { 1000, 1024, 1200, 1250, 1280, 1500, 1536, 1600, 1792, 1875, 1920, 2000,
                         2048, 2400, 2500, 2560, 3000, 3072, 3200, 3584, 3750, 3840, 4000, 4096 };
constexpr int menu_size  = sizeof(menu) / sizeof(menu[0]);
constexpr int lds_values = 16384; // complex f32 values per workgroup (128 KiB)
constexpr int pm         = 16;    // target points per thread and pass

struct radix_list
{
    int n     = 0;
    int r[24] = {};
};

constexpr radix_list make_radices(int N)
{
    radix_list L;
    int m = N, a = 0;
    while (m % 2 == 0)
    {
        m /= 2;
        ++a;
    }
    while (a >= 4)
    {
        L.r[L.n++] = 16;
        a -= 4;
    }
    if (a)
        L.r[L.n++] = 1 << a;
    for (int p : { 7, 5, 3 })
        while (m % p == 0)
        {
            L.r[L.n++] = p;
            m /= p;
        }
    return L;
}

template <int N>
struct mshape
{
    static constexpr radix_list RL = make_radices(N);
    static constexpr int NP        = RL.n;
    static constexpr int radix(int p) { return RL.r[p]; }
    static constexpr int ns(int p)
    {
        int s = 1;
        for (int q = 0; q < p; ++q)
            s *= RL.r[q];
        return s;
    }
    static constexpr int bf(int p) { return (pm + radix(p) - 1) / radix(p); } // butterflies per thread
    static constexpr int tpt()
    {
        int t = 1;
        for (int p = 0; p < NP; ++p)
        {
            const int nb = N / radix(p);
            const int v  = (nb + bf(p) - 1) / bf(p);
            t            = v > t ? v : t;
        }
        return t;
    }
    static constexpr int TPT = tpt();
    static constexpr int twoff(int p)
    {
        int o = 0;
        for (int q = 0; q < p; ++q)
            if (ns(q) > 1)
                o += ns(q) * radix(q);
        return o;
    }
    static constexpr int twcount = twoff(NP);
    static_assert(ns(NP) == N, "menu lengths factor into 2, 3, 5 and 7");
};

template <int N, int ROLE>
struct mgeom
{
    using S = mshape<N>;
    static constexpr int width()
    {
        if (ROLE == role_direct)
            return S::TPT >= 256 ? 1 : 256 / S::TPT;
        // row kernels: at most 512 threads (measured: two workgroups per CU without register spills beat
        // the wider runs of 1024 threads for the 4000-point rows; column kernels keep the wider runs)
        const int maxbs = ROLE == role_row ? 512 : 1024;
        int w           = 16;
        while (w > 1 && (w * N > lds_values || w * S::TPT > maxbs))
            w /= 2;
        return w;
    }
    static constexpr int W       = width();
    static constexpr int BS      = S::TPT * W;
    static constexpr int LD      = ROLE == role_row ? N + 1 : N;
    static constexpr int logical = ROLE == role_row ? W * LD : W * N;
    static constexpr size_t lds_bytes = (size_t)(logical + logical / 16 + 1) * sizeof(cx<float>);
    static_assert(BS <= 1024, "workgroup too large");
};

template <int ROLE, int W, int LD>
__device__ __forceinline__ int laddr(int tb, int i)
{
    const int x = ROLE == role_col ? i * W + tb : tb * LD + i;
    return x + (x >> 4);
}

__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

struct margs
{
    const cx<float>* in = nullptr;
    cx<float>* out      = nullptr;
    const cx<float>* tw = nullptr; // pass tables of this kernel's length
    const cx<float>* ta = nullptr; // four-step: W_n^(h*2^lbits)
    const cx<float>* tb = nullptr; // four-step: W_n^l, l < 2^lbits
    long long count     = 0;       // transforms
    long long n         = 0;       // four-step: full length n1*n2
    fastdiv odiv;                  // four-step: the other dimension (col: n1, row: n2)
    int other     = 1;
    int lbits     = 0;
    unsigned nblk = 0;
};

// ceil(log2(v)) of a positive compile-time value
__host__ __device__ constexpr int mlog2c(int v)
{
    int l = 0;
    while ((1 << l) < v)
        ++l;
    return l;
}

// element e of this thread's transform in HBM: base + e*stride (col: a column; direct, row: a row)
template <int N, bool INV, int ROLE, int P>
__device__ __forceinline__ void mpass(const margs& a, cx<float>* lds, int tb, int t, bool valid, long long base,
                                      long long stride, long long col)
{
    using S          = mshape<N>;
    using G          = mgeom<N, ROLE>;
    constexpr int R  = S::radix(P);
    constexpr int NS = S::ns(P);
    constexpr int M  = S::bf(P);
    constexpr int NB = N / R;
    constexpr int TP = S::TPT;
    cx<float> u[M][R];
#pragma unroll
    for (int m = 0; m < M; ++m)
    {
        const int j = t + m * TP;
        if (j < NB)
        {
#pragma unroll
            for (int r = 0; r < R; ++r)
            {
                const int e = j + r * NB;
                if constexpr (P == 0)
                    u[m][r] = valid ? a.in[base + (long long)e * stride] : cx<float>{ 0.f, 0.f };
                else
                    u[m][r] = lds[laddr<ROLE, G::W, G::LD>(tb, e)];
            }
        }
    }
    if constexpr (P > 0)
        lds_barrier(); // every read of the previous pass's output is done before this pass writes
#pragma unroll
    for (int m = 0; m < M; ++m)
    {
        const int j = t + m * TP;
        if (j < NB)
        {
            if constexpr (NS > 1)
            {
                const int kt       = j % NS;
                const cx<float>* w = a.tw + S::twoff(P) + kt * R;
                if constexpr (R >= 4)
                {
                    // W^(kt*r) from the row's entries at r = 2^b, one complex product per value
                    cx<float> p[R];
#pragma unroll
                    for (int b = 0; (1 << b) < R; ++b)
                        p[1 << b] = w[1 << b];
#pragma unroll
                    for (int r = 3; r < R; ++r)
                        if ((r & (r - 1)) != 0)
                            p[r] = cmul(p[r - (1 << (mlog2c(r + 1) - 1))], p[1 << (mlog2c(r + 1) - 1)]);
#pragma unroll
                    for (int r = 1; r < R; ++r)
                        u[m][r] = ctw<INV>(u[m][r], p[r]);
                }
                else
                {
#pragma unroll
                    for (int r = 1; r < R; ++r)
                        u[m][r] = ctw<INV>(u[m][r], w[r]);
                }
            }
            butterfly<R, INV>(u[m]);
        }
    }
    if constexpr (P + 1 < S::NP)
    {
        // output r of butterfly j goes to (j - j mod Ns)*R + j mod Ns + r*Ns
#pragma unroll
        for (int m = 0; m < M; ++m)
        {
            const int j = t + m * TP;
            if (j < NB)
            {
                const int kt = NS > 1 ? j % NS : 0;
                const int b  = (j - kt) * R + kt;
#pragma unroll
                for (int r = 0; r < R; ++r)
                    lds[laddr<ROLE, G::W, G::LD>(tb, b + r * NS)] = u[m][r];
            }
        }
        lds_barrier();
        mpass<N, INV, ROLE, P + 1>(a, lds, tb, t, valid, base, stride, col);
    }
    else
    {
        // last pass: Ns*R = N and j < Ns, so output r of butterfly j is element j + r*Ns
        // Column kernel: W_n^(col*e) for e = t + m*TPT + r*Ns is W^(col*t) * A_m * B_r with
        // A_m = W^(col*m*TPT) and B_r = W^(col*r*Ns), each built from table values W^(col*TPT*2^b) and
        // W^(col*Ns*2^b) with one complex product per value (instead of a two-level gather per element)
        const unsigned long long mask = (1ull << a.lbits) - 1;
        [[maybe_unused]] cx<float> A[M], B[R];
        if constexpr (ROLE == role_col)
        {
            auto table = [&](unsigned long long q) { return cmul(a.ta[q >> a.lbits], a.tb[q & mask]); };
            const unsigned long long nm = (unsigned long long)a.n;
            A[0] = table((unsigned long long)col * (unsigned long long)t);
            cx<float> pa[mlog2c(M) > 0 ? mlog2c(M) : 1];
#pragma unroll
            for (int b = 0; (1 << b) < M; ++b)
                pa[b] = table(((unsigned long long)col * (unsigned long long)(TP << b)) % nm);
#pragma unroll
            for (int m = 1; m < M; ++m)
            {
                const int hb = mlog2c(m + 1) - 1;
                A[m]         = cmul(A[m - (1 << hb)], pa[hb]);
            }
            cx<float> pb[mlog2c(R) > 0 ? mlog2c(R) : 1];
#pragma unroll
            for (int b = 0; (1 << b) < R; ++b)
                pb[b] = table(((unsigned long long)col * (unsigned long long)(NS << b)) % nm);
            B[0] = { 1.f, 0.f };
#pragma unroll
            for (int r = 1; r < R; ++r)
            {
                const int hb = mlog2c(r + 1) - 1;
                B[r]         = r == (1 << hb) ? pb[hb] : cmul(B[r - (1 << hb)], pb[hb]);
            }
        }
#pragma unroll
        for (int m = 0; m < M; ++m)
        {
            const int j = t + m * TP;
            if (j < NB)
            {
#pragma unroll
                for (int r = 0; r < R; ++r)
                {
                    const int e = j + r * NS;
                    if constexpr (ROLE == role_row)
                        lds[laddr<ROLE, G::W, G::LD>(tb, e)] = u[m][r];
                    else if constexpr (ROLE == role_col)
                    {
                        if (valid)
                        {
                            const cx<float> w = r == 0 ? A[m] : cmul(A[m], B[r]);
                            a.out[base + (long long)e * stride] = ctw<INV>(u[m][r], w);
                        }
                    }
                    else if (valid)
                        a.out[base + e] = u[m][r];
                }
            }
        }
    }
}

template <int N, bool INV, int ROLE>
__global__ void __launch_bounds__((mgeom<N, ROLE>::BS)) k_mix(margs a)
{
    using S          = mshape<N>;
    using G          = mgeom<N, ROLE>;
    constexpr int TP = S::TPT;
    constexpr int W  = G::W;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<float>* lds = reinterpret_cast<cx<float>*>(smem_raw);

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
    const int tb      = ROLE == role_col ? tid % W : tid / TP;
    const int t       = ROLE == role_col ? tid / W : tid % TP;
    const long long g = (long long)blk * W + tb;
    const bool valid  = g < a.count;
    long long base = 0, stride = 1, col = 0;
    if constexpr (ROLE == role_col)
    {
        const long long bl = valid ? (long long)fdiv((uint32_t)g, a.odiv) : 0;
        col                = g - bl * a.other;
        base               = bl * a.n + col;
        stride             = a.other;
    }
    else
        base = g * N;
    mpass<N, INV, ROLE, 0>(a, lds, tb, t, valid, base, stride, col);

    if constexpr (ROLE == role_row)
    {
        lds_barrier();
        // transposed store: row index fastest, so lanes write W consecutive outputs k2 + n2*k1
        const int tb2      = tid % W;
        const int i0       = tid / W;
        const long long g2 = (long long)blk * W + tb2;
        if (g2 < a.count)
        {
            const long long bl = (long long)fdiv((uint32_t)g2, a.odiv);
            const long long k2 = g2 - bl * a.other;
            const long long ob = bl * a.n + k2;
            for (int i = i0; i < N; i += TP)
                a.out[ob + (long long)i * a.other] = lds[laddr<ROLE, W, G::LD>(tb2, i)];
        }
    }
}

template <int N, bool INV, int ROLE>
void mlaunch(margs a, long long count, hipStream_t s)
{
    using G                  = mgeom<N, ROLE>;
    static const bool lds_ok = (allow_lds(k_mix<N, INV, ROLE>, G::lds_bytes), true);
    (void)lds_ok;
    if (count <= 0)
        return;
    a.count             = count;
    const long long nb  = (count + G::W - 1) / G::W;
    const long long grd = ROLE == role_direct ? nb : (nb + 7) / 8 * 8;
    if (grd > 0x7fffffffll)
        throw std::runtime_error("batch too large for one launch");
    a.nblk = (unsigned)nb;
    timing::span ts([] { return timing::kernel_name("k_mix", N, INV, ROLE); },
                    2 * (long long)sizeof(cx<float>) * count * N, s);
    hipLaunchKernelGGL((k_mix<N, INV, ROLE>), dim3((unsigned)grd), dim3(G::BS), G::lds_bytes, s, a);
    hip_check(hipGetLastError(), "mixed-radix FFT launch");
}

using launch_fn = void (*)(margs, long long, hipStream_t);

template <bool INV, int ROLE, size_t... I>
constexpr std::array<launch_fn, menu_size> make_launchers(std::index_sequence<I...>)
{
    return { &mlaunch<menu[I], INV, ROLE>... };
}
template <int ROLE, size_t... I>
constexpr std::array<int, menu_size> make_widths(std::index_sequence<I...>)
{
    return { mgeom<menu[I], ROLE>::W... };
}
template <size_t... I>
constexpr std::array<int, menu_size> make_twcounts(std::index_sequence<I...>)
{
    return { mshape<menu[I]>::twcount... };
}
template <size_t... I>
constexpr std::array<radix_list, menu_size> make_lists(std::index_sequence<I...>)
{
    return { mshape<menu[I]>::RL... };
}

using seq = std::make_index_sequence<menu_size>;

launch_fn launcher(int idx, bool inv, int role)
{
    static const std::array<launch_fn, menu_size> t[2][3] = {
        { make_launchers<false, role_direct>(seq{}), make_launchers<false, role_col>(seq{}),
          make_launchers<false, role_row>(seq{}) },
        { make_launchers<true, role_direct>(seq{}), make_launchers<true, role_col>(seq{}),
          make_launchers<true, role_row>(seq{}) }
    };
    return t[inv ? 1 : 0][role][idx];
}

constexpr std::array<int, menu_size> width_col = make_widths<role_col>(seq{});
constexpr std::array<int, menu_size> width_row = make_widths<role_row>(seq{});
constexpr std::array<int, menu_size> twcounts  = make_twcounts(seq{});
constexpr std::array<radix_list, menu_size> lists = make_lists(seq{});

int menu_index(size_t n)
{
    for (int i = 0; i < menu_size; ++i)
        if ((size_t)menu[i] == n)
            return i;
    return -1;
}

// pass tables of menu length idx: W_{Ns*R}^(r*k) at k*R + r for every pass with Ns > 1
void gen_tables(int idx, cx<float>* tw, hipStream_t s)
{
    const radix_list& L = lists[idx];
    long long ns = 1, off = 0;
    for (int p = 0; p < L.n; ++p)
    {
        const int r = L.r[p];
        if (ns > 1)
        {
            gen_twiddles(tw + off, ns * r, ns * r, 0, 0, tw_radix_full, r, s);
            off += ns * r;
        }
        ns *= r;
    }
}

// four-step split: both factors menu lengths, widest narrower run, then balance
bool split(size_t n, int& i2, int& i1)
{
    int best_w = -1;
    double best_bal = 1e300;
    i2 = i1 = -1;
    for (int a = 0; a < menu_size; ++a)
    {
        const size_t d = (size_t)menu[a];
        if (n % d)
            continue;
        const int b = menu_index(n / d);
        if (b < 0)
            continue;
        const int w      = std::min(width_col[a], width_row[b]);
        const double bal = std::fabs(std::log((double)d) - std::log((double)(n / d)));
        if (w > best_w || (w == best_w && bal < best_bal))
        {
            best_w   = w;
            best_bal = bal;
            i2       = a; // column length n2
            i1       = b; // row length n1
        }
    }
    return best_w > 0;
}

} // namespace mx

template <typename T>
struct mixed_fft<T>::impl
{
    size_t n  = 0;
    bool four = false;
    int idx = -1, i1 = -1, i2 = -1; // menu indices: direct length, or rows n1 and columns n2
    int lbits = 0;
    dev_mem tw;
    long long o_row = 0, o_ta = 0, o_tb = 0;
    dev_mem scratch;
    long long chunk = 1;
};

template <typename T>
bool mixed_fft<T>::supports(size_t n)
{
    if (sizeof(T) != 4 || n == 0 || (n & (n - 1)) == 0)
        return false;
    if (mx::menu_index(n) >= 0)
        return true;
    int a, b;
    return n <= ((size_t)1 << 28) && mx::split(n, a, b);
}

template <typename T>
mixed_fft<T>::mixed_fft(size_t n, hipStream_t s) : p(new impl)
{
    if (!supports(n))
        throw std::invalid_argument("mixed_fft: unsupported length");
    p->n   = n;
    p->idx = mx::menu_index(n);
    if (p->idx >= 0)
    {
        p->tw.reset((size_t)std::max(1, mx::twcounts[p->idx]) * sizeof(cx<float>));
        mx::gen_tables(p->idx, p->tw.template as<cx<float>>(), s);
        return;
    }
    p->four = true;
    mx::split(n, p->i2, p->i1);
    int lb = 0;
    while ((size_t(1) << lb) < n)
        ++lb;
    p->lbits            = (lb + 1) / 2;
    const long long nTB = 1ll << p->lbits;
    const long long nTA = (long long)(n >> p->lbits) + 1;
    p->o_row            = mx::twcounts[p->i2] + 1;
    p->o_ta             = p->o_row + mx::twcounts[p->i1] + 1;
    p->o_tb             = p->o_ta + nTA;
    p->tw.reset((size_t)(p->o_tb + nTB) * sizeof(cx<float>));
    cx<float>* t = p->tw.template as<cx<float>>();
    mx::gen_tables(p->i2, t, s);
    mx::gen_tables(p->i1, t + p->o_row, s);
    gen_twiddles(t + p->o_ta, nTA, (long long)n, nTB, 0, tw_linear, 0, s);
    gen_twiddles(t + p->o_tb, nTB, (long long)n, 1, 0, tw_linear, 0, s);
    size_t cap = (size_t)1 << 30; // rounds of at most 1 GiB, as the power-of-two four-step
    if (const char* e = std::getenv("KFR_AMD_FOURSTEP_CHUNK_MB"))
        cap = (size_t)std::max(1, atoi(e)) << 20;
    // fewer than 2^31 columns or rows per launch (32-bit indices)
    const long long wide = std::max(mx::menu[p->i1], mx::menu[p->i2]);
    const long long lim  = std::max<long long>(1, ((1ll << 31) - 1) / wide);
    p->chunk             = std::min<long long>((long long)std::max<size_t>(1, cap / (n * sizeof(cx<float>))), lim);
}

template <typename T>
mixed_fft<T>::~mixed_fft() = default;

template <typename T>
void mixed_fft<T>::run(bool inverse, const void* in, void* out, long long batch, hipStream_t s, workspace w)
{
    if (batch <= 0)
        return;
    const cx<float>* x = static_cast<const cx<float>*>(in);
    cx<float>* y       = static_cast<cx<float>*>(out);
    const cx<float>* t = p->tw.template as<cx<float>>();
    if (!p->four)
    {
        mx::margs a;
        a.in  = x;
        a.out = y;
        a.tw  = t;
        mx::launcher(p->idx, inverse, role_direct)(a, batch, s);
        return;
    }
    const long long n  = (long long)p->n;
    const long long n1 = mx::menu[p->i1], n2 = mx::menu[p->i2];
    const long long chunk = rounds_in(w, p->chunk, (size_t)n * sizeof(cx<float>), 1);
    cx<float>* z;
    if (w.p)
        z = static_cast<cx<float>*>(w.p);
    else
    {
        p->scratch.ensure((size_t)std::min(chunk, batch) * n * sizeof(cx<float>));
        z = p->scratch.template as<cx<float>>();
    }
    for (long long b0 = 0; b0 < batch; b0 += chunk)
    {
        const long long nb = std::min(chunk, batch - b0);
        mx::margs a;
        a.ta    = t + p->o_ta;
        a.tb    = t + p->o_tb;
        a.lbits = p->lbits;
        a.n     = n;
        a.in    = x + b0 * n;
        a.out   = z;
        a.tw    = t;
        a.other = (int)n1;
        a.odiv  = make_fastdiv((uint32_t)n1);
        mx::launcher(p->i2, inverse, role_col)(a, nb * n1, s);
        a.in    = z;
        a.out   = y + b0 * n;
        a.tw    = t + p->o_row;
        a.other = (int)n2;
        a.odiv  = make_fastdiv((uint32_t)n2);
        mx::launcher(p->i1, inverse, role_row)(a, nb * n2, s);
    }
}

template <typename T>
std::string mixed_fft<T>::describe() const
{
    auto radices = [](int idx)
    {
        std::string r;
        for (int q = 0; q < mx::lists[idx].n; ++q)
            r += (q ? "x" : "") + std::to_string(mx::lists[idx].r[q]);
        return r;
    };
    if (!p->four)
        return "mixed radices=" + radices(p->idx);
    return "mixed four-step n1=" + std::to_string(mx::menu[p->i1]) + " n2=" + std::to_string(mx::menu[p->i2]) +
           " radices=" + radices(p->i2) + "|" + radices(p->i1) + " columns/workgroup=" +
           std::to_string(mx::width_col[p->i2]) + " rows/workgroup=" + std::to_string(mx::width_row[p->i1]);
}

template class mixed_fft<float>;
template class mixed_fft<double>;

} // namespace kfr_amd
