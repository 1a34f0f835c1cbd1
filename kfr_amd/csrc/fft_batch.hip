// Uneven batches in one launch: a persistent kernel pulls work items (groups of transforms of one
// length and direction) from a device queue, longest first, and runs each item's transforms in LDS
// with the schedule of its length. Every length's Stockham geometry, twiddles and (for lengths with
// a prime factor of 101 or more, as the reference's planner decides) Bluestein tables sit in device
// memory, so a single kernel covers any mix of lengths. Inverse transforms are computed as
// conj(DFT(conj x)), so the kernel carries one set of butterflies. The queue reset and the kernel
// are captured once into a hipGraph and replayed.
//
// An autotuner times candidate schedules of each length on the device: the order of its radix passes,
// the LDS padding of its tile, and the transforms per item (without tuning: the planner's order, one
// padding slot per 16 values, and as many transforms as the occupancy limits allow).
#include "fft_batch.hpp"

#include "fft_core.hpp"
#include "plan_desc.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <stdexcept>
#include <utility>
#include <vector>

namespace kfr_amd
{
namespace bt
{

constexpr int block          = max_block_mixed;  // threads per workgroup
constexpr size_t lds_cap     = 78 * 1024;         // LDS per workgroup: two workgroups per CU
constexpr int max_candidates = 5;

struct len_desc
{
    lds_geom G;          // geometry of the transform held in LDS (n points, or M for Bluestein)
    long long tw    = 0; // its Stockham twiddles (offset into the table, complex values)
    long long chirp = 0; // Bluestein chirp (n values)
    long long H     = 0; // Bluestein kernel spectrum (M values)
    int n           = 0;
    int bluestein   = 0;
};

struct work_item
{
    int len;   // index into the length descriptors
    int first; // first entry in the member list
    int count; // transforms in this item (at most the length's transforms per item)
};

// A member entry is (transform index << 1) | inverse: the transforms of one length share items
// whatever their direction (fewer part-filled items than grouping by length and direction).
constexpr size_t max_transforms = (size_t)1 << 30;

__global__ void k_queue_reset(unsigned* counter)
{
    if (threadIdx.x == 0)
        counter[0] = 0u;
}

// at most 128 VGPRs (4 waves per SIMD), so two workgroups share a CU: measured faster than the
// unconstrained build (185 VGPRs, one workgroup per CU) despite a little spilling
template <typename T>
__global__ void __launch_bounds__(block) __attribute__((amdgpu_waves_per_eu(4))) k_batch(const len_desc* __restrict__ lens, const work_item* __restrict__ items,
                                                 int nitems, const int* __restrict__ members,
                                                 const long long* __restrict__ offs, const cx<T>* __restrict__ tab,
                                                 const cx<T>* in, cx<T>* out, unsigned* counter)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    // the next item is claimed when the current one starts, and its index published when the current
    // one ends, so the atomic's round trip overlaps the transform instead of preceding it
    __shared__ int s_item;
    if (threadIdx.x == 0)
        s_item = (int)atomicAdd(counter, 1u);
    __syncthreads();
    for (;;)
    {
        const int it = s_item;
        __syncthreads(); // every thread has read s_item before thread 0 overwrites it below
        if (it >= nitems)
            return; // uniform across the workgroup
        unsigned next = 0;
        if (threadIdx.x == 0)
            next = atomicAdd(counter, 1u);
        const work_item I = items[it];
        const len_desc& D = lens[I.len];
        const lds_geom& G = D.G;
        const int tid     = threadIdx.x;
        const int tb      = tid / G.tpt;
        const bool active = tb < I.count;
        fft_ctx<T> c;
        c.lds   = lds;
        c.tb    = tb;
        c.base  = active ? tb * G.ld : 0;
        c.ls    = 1;
        c.ps    = G.padshift;
        c.t     = active ? tid - tb * G.tpt : G.n; // idle threads skip every element loop
        c.valid = active;
        c.g     = 0;
        const int mbr       = active ? members[I.first + tb] : 0;
        const long long off = active ? offs[mbr >> 1] : 0;
        const bool inv      = (mbr & 1) != 0;
        const int n         = D.n;
        const cx<T>* ch     = tab + D.chirp;

        // load; the inverse runs as conj(DFT(conj x)); Bluestein premultiplies by the chirp and pads.
        // A thread holds at most EM values of one transform (every pass gives it at most
        // ceil(pm/R)*R <= pm + 6 of them): the loads of all of them are issued before the first is
        // used, so their latencies overlap instead of adding up; any value beyond EM (none for the
        // schedules the planner makes) takes the plain loop.
        constexpr int EM = pmax_of<T> + 7;
        {
            cx<T> v[EM];
#pragma unroll
            for (int k = 0; k < EM; ++k)
            {
                const int i = c.t + k * G.tpt;
                v[k]        = { T(0), T(0) };
                if (i < n)
                    v[k] = in[off + i];
            }
#pragma unroll
            for (int k = 0; k < EM; ++k)
            {
                const int i = c.t + k * G.tpt;
                if (i < G.n)
                {
                    cx<T> x = v[k];
                    if (i < n)
                    {
                        if (inv)
                            x = cconj(x);
                        if (D.bluestein)
                            x = cmul(x, ch[i]);
                    }
                    c.at(i) = x;
                }
            }
        }
        for (int i = c.t + EM * G.tpt; i < G.n; i += G.tpt)
        {
            cx<T> v = { T(0), T(0) };
            if (i < n)
            {
                v = in[off + i];
                if (inv)
                    v = cconj(v);
                if (D.bluestein)
                    v = cmul(v, ch[i]);
            }
            c.at(i) = v;
        }
        __syncthreads();
        fft_passes<T, false, true, true>(G, tab + D.tw, c);
        if (D.bluestein)
        {
            // IFFT_M(A conj(H)) = conj(FFT_M(conj(A) H)); H carries the 1/M
            const cx<T>* H = tab + D.H;
            for (int i = c.t; i < G.n; i += G.tpt)
                c.at(i) = cmul(cconj(c.at(i)), H[i]);
            __syncthreads();
            fft_passes<T, false, true, true>(G, tab + D.tw, c);
        }
        for (int i = c.t; i < n; i += G.tpt)
        {
            cx<T> v = c.at(i);
            if (D.bluestein)
                v = cmul(cconj(v), ch[i]);
            v = inv ? cconj(v) : v;
            // nontemporal: the outputs are not read again by this launch (measured 4% faster)
            using V2 = T __attribute__((ext_vector_type(2)));
            __builtin_nontemporal_store(V2{ v.re, v.im }, reinterpret_cast<V2*>(out + off + i));
        }
        if (threadIdx.x == 0)
            s_item = (int)next;
        __syncthreads(); // LDS is rewritten by the next item, whose index is now in s_item
    }
}

// host-side description of one distinct length
template <typename T>
struct host_len
{
    size_t n  = 0;
    size_t m  = 0; // LDS transform length (n, or the Bluestein M)
    bool bs   = false;
    stockham S;                            // schedule of the chosen pass order
    std::vector<std::vector<int>> orders;  // candidate pass orders (the planner's first)
    int order    = 0;                      // chosen pass order
    int padshift = 4;                      // chosen LDS padding: one slot per 2^padshift values
    int tpb_max  = 1;
    int tpb      = 1;
    long long tw = 0, chirp = 0, H = 0;
};

// LDS paddings tried: one slot per 16 or per 32 values, or none (the tile is sized for one per 16)
constexpr int padshifts[] = { 4, 5, 30 };

// Pass orders tried for one length: the planner's (powers of two, largest first, then 7, 5, 3 and the
// direct primes), its reverse, and the odd factors ahead of the powers of two. Every order of one
// radix multiset has the same threads per transform and table size; each pass reads its own stride
// and twiddles from the length's descriptor, so any order computes the same transform.
inline std::vector<std::vector<int>> radix_orders(const std::vector<int>& r)
{
    std::vector<std::vector<int>> out{ r };
    std::vector<int> rev(r.rbegin(), r.rend());
    std::vector<int> odd_first;
    for (int x : r)
        if (x % 2 != 0)
            odd_first.push_back(x);
    for (int x : r)
        if (x % 2 == 0)
            odd_first.push_back(x);
    for (auto& o : { rev, odd_first })
        if (std::find(out.begin(), out.end(), o) == out.end())
            out.push_back(o);
    return out;
}

template <typename T>
void set_order(host_len<T>& h, int k)
{
    h.order = k;
    if (!stockham_from_radices((long long)h.m, h.orders[k], h.S, pmax_of<T>, true))
        throw std::logic_error("batch: pass order does not factor its length");
}

// forced choices for tests and measurements: KFR_AMD_BATCH_ORDER (index of the pass order, modulo the
// number of candidates) and KFR_AMD_BATCH_PADSHIFT (4, 5 or 30); -1 when unset
inline int env_choice(const char* name)
{
    const char* e = std::getenv(name);
    return e ? std::atoi(e) : -1;
}

inline size_t align2(size_t v) { return (v + 1) / 2 * 2; }

template <typename T>
size_t lds_of(size_t m, int tpb)
{
    return lds_padded_elems((size_t)tpb * m) * sizeof(cx<T>);
}

inline int cu_count()
{
    static const int cus = []
    {
        int dev = 0, n = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hip_check(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
        return std::max(1, n);
    }();
    return cus;
}

// the kernel may use up to lds_cap of dynamic LDS; raised once, before any launch or graph capture
template <typename T>
void allow_batch_lds()
{
    static const bool done = (allow_lds(k_batch<T>, lds_cap), true);
    (void)done;
}

template <typename T>
unsigned grid_for(size_t lds_bytes)
{
    allow_batch_lds<T>();
    int occ = 0;
    hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_batch<T>, block, lds_bytes), "occupancy query");
    return (unsigned)std::max(1, occ) * (unsigned)cu_count();
}

template <typename T>
void launch(const dev_mem& lens, const dev_mem& items, int nitems, const dev_mem& members, const dev_mem& offs,
            const dev_mem& tab, const void* in, void* out, const dev_mem& counter, size_t lds_bytes, unsigned grid,
            hipStream_t s)
{
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, s, counter.as<unsigned>());
    hip_check(hipGetLastError(), "queue reset launch");
    hipLaunchKernelGGL(k_batch<T>, dim3(grid), dim3(block), lds_bytes, s, lens.as<len_desc>(), items.as<work_item>(),
                       nitems, members.as<int>(), offs.as<long long>(), tab.as<cx<T>>(),
                       static_cast<const cx<T>*>(in), static_cast<cx<T>*>(out), counter.as<unsigned>());
    hip_check(hipGetLastError(), "batch kernel launch");
}

template <typename T>
len_desc make_desc(const host_len<T>& h, int tpb)
{
    len_desc d;
    check_pass_radices<T>(h.S, true);
    d.G          = make_geom(h.S, tpb, 0, 1, (int)h.m);
    d.G.padshift = h.padshift;
    d.tw        = h.tw;
    d.chirp     = h.chirp;
    d.H         = h.H;
    d.n         = (int)h.n;
    d.bluestein = h.bs ? 1 : 0;
    return d;
}

struct tuning
{
    int order = 0, padshift = 4, tpb = 1;
};

// process-wide cache of tuned schedules, keyed by (element size, length)
inline std::map<std::pair<int, size_t>, tuning>& tuned()
{
    static std::map<std::pair<int, size_t>, tuning> m;
    return m;
}
inline std::mutex& tuned_mu()
{
    static std::mutex mu;
    return mu;
}

// Times candidate schedules of one length on synthetic data and leaves the fastest per transform in h:
// first every pass order with every LDS padding at the largest transforms-per-item, then the
// transforms per item (the largest and up to four halvings of it, and 1) with the winning order and
// padding. Each candidate runs one untimed launch and three timed ones (median). Forced choices
// (KFR_AMD_BATCH_ORDER, KFR_AMD_BATCH_PADSHIFT) are kept and not searched.
template <typename T>
void autotune(host_len<T>& h, int force_order, int force_ps, hipStream_t s)
{
    std::vector<int> orders, pads, tpbs;
    for (int k = 0; k < (int)h.orders.size(); ++k)
        if (force_order < 0 || k == force_order % (int)h.orders.size())
            orders.push_back(k);
    for (int ps : padshifts)
        if (force_ps < 0 || ps == force_ps)
            pads.push_back(ps);
    if (pads.empty())
        pads.push_back(4);
    for (int k = 0; k < max_candidates - 1; ++k)
        if ((h.tpb_max >> k) >= 1)
            tpbs.push_back(h.tpb_max >> k);
    tpbs.push_back(1);
    std::sort(tpbs.begin(), tpbs.end());
    tpbs.erase(std::unique(tpbs.begin(), tpbs.end()), tpbs.end());

    // one table for the candidates: Bluestein chirp and spectrum (order-independent), then the
    // twiddles of the order being timed
    const size_t c0 = 0, h0 = h.bs ? align2(h.n) : 0, t0 = h.bs ? h0 + align2(h.m) : 0;
    dev_mem tab((t0 + align2((size_t)std::max<long long>(1, h.S.tw_count))) * sizeof(cx<T>));
    if (h.bs)
        bluestein_tables<T>(h.n, h.m, tab.as<cx<T>>() + c0, tab.as<cx<T>>() + h0, s);
    hipEvent_t ev[4];
    for (auto& e : ev)
        hip_check(hipEventCreate(&e), "event");
    dev_mem lens(sizeof(len_desc)), counter(sizeof(unsigned)), items, members, offs, data;
    int loaded_order = -1;
    auto time_one = [&](int order, int ps, int tpb) -> double
    {
        if (order != loaded_order)
        {
            set_order(h, order);
            gen_stockham_twiddles(h.S, tab.as<cx<T>>() + t0, s);
            loaded_order = order;
        }
        h.padshift          = ps;
        const size_t lb     = lds_of<T>(h.m, tpb);
        const unsigned grid = grid_for<T>(lb);
        const int nitems    = (int)(2 * grid);
        const long long nt  = (long long)nitems * tpb;
        std::vector<work_item> it(nitems);
        std::vector<int> mem(nt);
        std::vector<long long> of(nt);
        for (int i = 0; i < nitems; ++i)
            it[i] = { 0, i * tpb, tpb };
        for (long long j = 0; j < nt; ++j)
        {
            mem[j] = (int)(j << 1);
            of[j]  = j * (long long)h.n;
        }
        host_len<T> hh = h;
        hh.tw          = (long long)t0;
        hh.chirp       = (long long)c0;
        hh.H           = (long long)h0;
        const len_desc d = make_desc(hh, tpb);
        items.ensure(it.size() * sizeof(work_item));
        members.ensure(mem.size() * sizeof(int));
        offs.ensure(of.size() * sizeof(long long));
        data.ensure((size_t)nt * h.n * sizeof(cx<T>));
        hip_check(hipMemcpyAsync(lens.ptr, &d, sizeof(d), hipMemcpyHostToDevice, s), "tuning upload");
        hip_check(hipMemcpyAsync(items.ptr, it.data(), it.size() * sizeof(work_item), hipMemcpyHostToDevice, s),
                  "tuning upload");
        hip_check(hipMemcpyAsync(members.ptr, mem.data(), mem.size() * sizeof(int), hipMemcpyHostToDevice, s),
                  "tuning upload");
        hip_check(hipMemcpyAsync(offs.ptr, of.data(), of.size() * sizeof(long long), hipMemcpyHostToDevice, s),
                  "tuning upload");
        hip_check(hipMemsetAsync(data.ptr, 0, (size_t)nt * h.n * sizeof(cx<T>), s), "tuning data");
        launch<T>(lens, items, nitems, members, offs, tab, data.ptr, data.ptr, counter, lb, grid, s); // warm
        hip_check(hipEventRecord(ev[0], s), "event");
        for (int r = 0; r < 3; ++r)
        {
            launch<T>(lens, items, nitems, members, offs, tab, data.ptr, data.ptr, counter, lb, grid, s);
            hip_check(hipEventRecord(ev[r + 1], s), "event");
        }
        hip_check(hipEventSynchronize(ev[3]), "tuning");
        float ms[3];
        for (int r = 0; r < 3; ++r)
            hip_check(hipEventElapsedTime(&ms[r], ev[r], ev[r + 1]), "event");
        std::sort(ms, ms + 3);
        return (double)ms[1] / (double)nt;
    };
    try
    {
        double best = 1e300;
        tuning b{ orders[0], pads[0], h.tpb_max };
        for (int o : orders)
            for (int ps : pads)
            {
                const double t = time_one(o, ps, h.tpb_max);
                if (t < best)
                {
                    best = t;
                    b    = { o, ps, h.tpb_max };
                }
            }
        for (int tpb : tpbs)
        {
            if (tpb == h.tpb_max)
                continue;
            const double t = time_one(b.order, b.padshift, tpb);
            if (t < best)
            {
                best  = t;
                b.tpb = tpb;
            }
        }
        hip_check(hipStreamSynchronize(s), "tuning");
        set_order(h, b.order);
        h.padshift = b.padshift;
        h.tpb      = b.tpb;
    }
    catch (...)
    {
        for (auto& e : ev)
            (void)hipEventDestroy(e);
        throw;
    }
    for (auto& e : ev)
        (void)hipEventDestroy(e);
}

} // namespace bt

template <typename T>
struct gpu_batch<T>::impl
{
    size_t count = 0, total = 0;
    dev_mem lens, items, members, offs, tab, counter;
    int nitems       = 0;
    size_t lds_bytes = 0;
    unsigned grid    = 0;
    std::string info;
    hipStream_t cap       = nullptr;
    hipGraph_t graph      = nullptr;
    hipGraphExec_t exec   = nullptr;
    const void* graph_in  = nullptr;
    void* graph_out       = nullptr;

    ~impl()
    {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        if (cap)
            (void)hipStreamDestroy(cap);
    }
};

template <typename T>
gpu_batch<T>::gpu_batch(size_t count, const size_t* sizes, const int* inverse, bool autotune) : p(new impl)
{
    using namespace bt;
    if (count == 0)
        throw std::invalid_argument("empty batch");
    if (count >= max_transforms)
        throw std::invalid_argument("batch has too many transforms");
    p->count          = count;
    hipStream_t s     = library_stream();
    const dtype_t dt  = sizeof(T) == 4 ? dtype_t::f32 : dtype_t::f64;
    // distinct lengths
    std::map<size_t, int> index;
    std::vector<host_len<T>> L;
    std::vector<long long> of(count);
    long long pos = 0;
    for (size_t i = 0; i < count; ++i)
    {
        const size_t n = sizes[i];
        if (n == 0)
            throw std::invalid_argument("transform length must be positive");
        of[i] = pos;
        pos += (long long)n;
        if (index.count(n))
            continue;
        host_len<T> h;
        h.n  = n;
        h.bs = describe_dft(dt, n, false).arblen;
        h.m  = h.bs ? next_pow2(2 * n - 1) : n;
        if (!make_stockham((long long)h.m, h.S, pmax_of<T>, true) || h.S.tpt > block || lds_of<T>(h.m, 1) > lds_cap)
            throw std::invalid_argument("transform length " + std::to_string(n) +
                                        " does not fit one workgroup of the batch kernel");
        h.orders = radix_orders(h.S.radix);
        h.tpb_max = std::max(1, block / h.S.tpt);
        while (h.tpb_max > 1 && lds_of<T>(h.m, h.tpb_max) > lds_cap)
            --h.tpb_max;
        index[n] = (int)L.size();
        L.push_back(h);
    }
    p->total = (size_t)pos;
    // schedules: tuned (cached per process), forced, or the planner's order, padding of one slot per 16
    // values and the largest transforms-per-item that fits
    const int force_order = env_choice("KFR_AMD_BATCH_ORDER");
    const int force_ps    = env_choice("KFR_AMD_BATCH_PADSHIFT");
    if (force_ps >= 0 && std::find(std::begin(padshifts), std::end(padshifts), force_ps) == std::end(padshifts))
        throw std::invalid_argument("KFR_AMD_BATCH_PADSHIFT must be 4, 5 or 30");
    int tuned_count = 0;
    for (auto& h : L)
    {
        h.tpb = h.tpb_max;
        if (force_order >= 0)
            set_order(h, force_order % (int)h.orders.size());
        if (force_ps >= 0)
            h.padshift = force_ps;
        if (!autotune)
            continue;
        const std::pair<int, size_t> key{ (int)sizeof(T), h.n };
        if (force_order < 0 && force_ps < 0)
        {
            std::lock_guard<std::mutex> lock(tuned_mu());
            auto f = tuned().find(key);
            if (f != tuned().end())
            {
                set_order(h, f->second.order);
                h.padshift = f->second.padshift;
                h.tpb      = std::min(f->second.tpb, h.tpb_max);
                continue;
            }
        }
        bt::autotune<T>(h, force_order, force_ps, s);
        ++tuned_count;
        if (force_order < 0 && force_ps < 0)
        {
            std::lock_guard<std::mutex> lock(tuned_mu());
            tuned()[key] = { h.order, h.padshift, h.tpb };
        }
    }
    // tables: per length its twiddles, and for Bluestein its chirp and spectrum
    size_t tabsz = 0;
    for (auto& h : L)
    {
        h.tw = (long long)tabsz;
        tabsz += align2((size_t)std::max<long long>(1, h.S.tw_count));
        if (h.bs)
        {
            h.chirp = (long long)tabsz;
            tabsz += align2(h.n);
            h.H = (long long)tabsz;
            tabsz += align2(h.m);
        }
    }
    p->tab.reset(tabsz * sizeof(cx<T>));
    cx<T>* tb = p->tab.template as<cx<T>>();
    for (auto& h : L)
    {
        gen_stockham_twiddles(h.S, tb + h.tw, s);
        if (h.bs)
            bluestein_tables<T>(h.n, h.m, tb + h.chirp, tb + h.H, s);
    }
    hip_check(hipStreamSynchronize(s), "batch tables");
    // group transforms by length (each member carries its own direction) and cut the groups into
    // items, costliest first
    std::map<int, std::vector<int>> groups;
    for (size_t i = 0; i < count; ++i)
        groups[index[sizes[i]]].push_back((int)(i << 1) | (inverse && inverse[i] ? 1 : 0));
    std::vector<int> mem;
    mem.reserve(count);
    std::vector<std::pair<double, work_item>> its;
    for (auto& [key, ids] : groups)
    {
        const host_len<T>& h = L[key];
        const double unit    = (double)h.m * std::log2((double)std::max<size_t>(2, h.m)) * (h.bs ? 2.0 : 1.0);
        for (size_t a = 0; a < ids.size(); a += h.tpb)
        {
            const int cnt = (int)std::min<size_t>(h.tpb, ids.size() - a);
            work_item w{ key, (int)mem.size(), cnt };
            for (int j = 0; j < cnt; ++j)
                mem.push_back(ids[a + j]);
            its.push_back({ unit * cnt, w });
        }
    }
    std::stable_sort(its.begin(), its.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
    std::vector<work_item> items;
    items.reserve(its.size());
    for (auto& w : its)
        items.push_back(w.second);
    p->nitems = (int)items.size();
    std::vector<len_desc> ds;
    for (auto& h : L)
    {
        ds.push_back(make_desc(h, h.tpb));
        p->lds_bytes = std::max(p->lds_bytes, lds_of<T>(h.m, h.tpb));
    }
    p->lens.reset(ds.size() * sizeof(len_desc));
    p->items.reset(items.size() * sizeof(work_item));
    p->members.reset(mem.size() * sizeof(int));
    p->offs.reset(of.size() * sizeof(long long));
    p->counter.reset(sizeof(unsigned));
    hip_check(hipMemcpyAsync(p->lens.ptr, ds.data(), ds.size() * sizeof(len_desc), hipMemcpyHostToDevice, s),
              "batch upload");
    hip_check(hipMemcpyAsync(p->items.ptr, items.data(), items.size() * sizeof(work_item), hipMemcpyHostToDevice, s),
              "batch upload");
    hip_check(hipMemcpyAsync(p->members.ptr, mem.data(), mem.size() * sizeof(int), hipMemcpyHostToDevice, s),
              "batch upload");
    hip_check(hipMemcpyAsync(p->offs.ptr, of.data(), of.size() * sizeof(long long), hipMemcpyHostToDevice, s),
              "batch upload");
    hip_check(hipStreamSynchronize(s), "batch upload");
    p->grid = std::min<unsigned>(grid_for<T>(p->lds_bytes), (unsigned)p->nitems);
    p->info = "gpu(gfx950): batch transforms=" + std::to_string(count) + " lengths=" + std::to_string(L.size()) +
              " items=" + std::to_string(p->nitems) + " workgroups=" + std::to_string(p->grid) +
              " lds=" + std::to_string(p->lds_bytes) + (autotune ? " autotuned=" + std::to_string(tuned_count) : "");
    // schedule choices over the distinct lengths: pass order (planner's, reversed, odd factors first),
    // LDS padding (1/16, 1/32, none)
    int by_order[3] = { 0, 0, 0 }, by_pad[3] = { 0, 0, 0 };
    for (auto& h : L)
    {
        ++by_order[std::min(h.order, 2)];
        ++by_pad[h.padshift == 4 ? 0 : h.padshift == 5 ? 1 : 2];
    }
    p->info += " orders=" + std::to_string(by_order[0]) + "/" + std::to_string(by_order[1]) + "/" +
               std::to_string(by_order[2]) + " padding=" + std::to_string(by_pad[0]) + "/" +
               std::to_string(by_pad[1]) + "/" + std::to_string(by_pad[2]);
}

template <typename T>
gpu_batch<T>::~gpu_batch() = default;

template <typename T>
size_t gpu_batch<T>::count() const
{
    return p->count;
}

template <typename T>
size_t gpu_batch<T>::total() const
{
    return p->total;
}

template <typename T>
std::string gpu_batch<T>::describe() const
{
    return p->info;
}

template <typename T>
void gpu_batch<T>::execute(void* out, const void* in, hipStream_t s)
{
    impl& q = *p;
    if (!q.exec || q.graph_in != in || q.graph_out != out)
    {
        if (q.exec)
        {
            (void)hipGraphExecDestroy(q.exec);
            q.exec = nullptr;
        }
        if (q.graph)
        {
            (void)hipGraphDestroy(q.graph);
            q.graph = nullptr;
        }
        if (!q.cap)
            hip_check(hipStreamCreateWithFlags(&q.cap, hipStreamNonBlocking), "capture stream");
        hip_check(hipStreamBeginCapture(q.cap, hipStreamCaptureModeThreadLocal), "graph capture");
        try
        {
            bt::launch<T>(q.lens, q.items, q.nitems, q.members, q.offs, q.tab, in, out, q.counter, q.lds_bytes, q.grid,
                          q.cap);
        }
        catch (...)
        {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(q.cap, &g);
            if (g)
                (void)hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(q.cap, &q.graph), "graph capture");
        hip_check(hipGraphInstantiate(&q.exec, q.graph, nullptr, nullptr, 0), "graph instantiation");
        q.graph_in  = in;
        q.graph_out = out;
    }
    hip_check(hipGraphLaunch(q.exec, s), "graph launch");
}

template class gpu_batch<float>;
template class gpu_batch<double>;

} // namespace kfr_amd
