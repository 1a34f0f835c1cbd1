// C ABI of the gfx950 DFT engine: the reference's kfr_dft_* / kfr_dft_real_* / kfr_dct_* /
// kfr_filter_* entry points (src/capi/capi.cpp) over GPU plans, plus the device extensions.
#include "../../include/kfr_amd_capi.h"
#include "fft_batch.hpp"
#include "fft_engine.hpp"
#include "fft_md.hpp"
#include "fft_pow2.hpp"
#include "fft_timing.hpp"
#include "plan_desc.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <map>
#include <tuple>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

using namespace kfr_amd;

namespace
{

// thread-local error text, cleared by every successful call (src/capi/capi.cpp:36-91)
thread_local std::array<char, 256> latest_error{};

void reset_error() { latest_error.fill(0); }
void set_error(const std::string& s)
{
    const size_t n = std::min(s.size(), latest_error.size() - 1);
    std::copy_n(s.begin(), n, latest_error.begin());
    std::fill(latest_error.begin() + n, latest_error.end(), 0);
}

struct timing_result
{
    std::string name;
    long long bytes;
    double ms;
};
std::vector<timing_result>& timing_results()
{
    static std::vector<timing_result> r;
    return r;
}

template <typename Fn, typename R>
R try_fn(Fn&& fn, R fallback)
{
    try
    {
        R r = fn();
        reset_error();
        return r;
    }
    catch (const std::exception& e)
    {
        set_error(e.what());
        return fallback;
    }
    catch (...)
    {
        set_error("(unknown exception)");
        return fallback;
    }
}
template <typename Fn>
void try_void(Fn&& fn)
{
    try_fn(
        [&]
        {
            fn();
            return 0;
        },
        0);
}

void check_plan(const void* p)
{
    if (!p)
        throw std::invalid_argument("plan is NULL");
}

bool is_device_ptr(const void* p)
{
    if (!p)
        return false;
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess)
    {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged || a.type == hipMemoryTypeUnified;
}

// NULL is the HIP default (null) stream, as in HIP and PyTorch (whose default stream handle is 0)
hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

// Orders the uses of one plan's device scratch (four-step intermediates, real and DCT half spectra,
// Bluestein buffers) across streams: work enqueued on a stream other than the one that used the
// plan last first waits for that stream's last work on this plan. Enqueues from several host
// threads are serialised by the mutex, which also guards the lazily grown scratch. Inside a stream
// capture nothing is recorded (the captured stream orders its own work).
struct stream_order
{
    std::mutex mu;
    hipEvent_t ev    = nullptr;
    hipStream_t last = nullptr;
    bool used        = false;
    stream_order()   = default;
    stream_order(const stream_order&) = delete;
    ~stream_order()
    {
        if (ev)
            (void)hipEventDestroy(ev);
    }
    template <typename Fn>
    void run(hipStream_t s, Fn&& fn)
    {
        std::lock_guard<std::mutex> lock(mu);
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        hip_check(hipStreamIsCapturing(s, &cs), "stream capture status");
        const bool capturing = cs != hipStreamCaptureStatusNone;
        if (!capturing && used && last != s)
            hip_check(hipStreamWaitEvent(s, ev, 0), "stream ordering");
        fn();
        if (!capturing)
        {
            if (!ev)
                hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
            hip_check(hipEventRecord(ev, s), "stream ordering");
            last = s;
            used = true;
        }
    }
};

// Runs fn(dev_out, dev_in) on device buffers: the user's pointers when both are device memory,
// otherwise one staging buffer (the transforms are safe in place for a single item) with copies.
struct stager
{
    dev_mem buf;
    std::mutex mu;
    template <typename Fn>
    void run(void* out, size_t out_bytes, const void* in, size_t in_bytes, Fn&& fn)
    {
        std::lock_guard<std::mutex> lock(mu);
        hipStream_t s = library_stream();
        if (is_device_ptr(in) && is_device_ptr(out))
        {
            // the reference's execute is synchronous: device inputs may still be in production on any
            // stream (the library's stream does not wait for the null stream), so wait for the device
            hip_check(hipDeviceSynchronize(), "execute");
            fn(out, in, s);
            hip_check(hipStreamSynchronize(s), "execute");
            return;
        }
        buf.ensure(std::max(in_bytes, out_bytes));
        hip_check(hipMemcpyAsync(buf.ptr, in, in_bytes, hipMemcpyDefault, s), "input copy");
        fn(buf.ptr, buf.ptr, s);
        hip_check(hipMemcpyAsync(out, buf.ptr, out_bytes, hipMemcpyDefault, s), "output copy");
        hip_check(hipStreamSynchronize(s), "execute");
    }
};

// Plans are reference counted: the creator's handle holds one reference, the plan cache one per
// cached plan, and a progressive execution one for its plan; kfr_*_delete_plan_* drops a reference.
struct plan_ref
{
    std::atomic<int> refs{ 1 };
    virtual ~plan_ref() = default;
};
void retain(plan_ref* p) { p->refs.fetch_add(1, std::memory_order_relaxed); }
void release(plan_ref* p)
{
    if (p && p->refs.fetch_sub(1, std::memory_order_acq_rel) == 1)
        delete p;
}

// a complex plan: one axis (dft_plan) or several (dft_plan_md), as the reference's C API keeps both
// behind the same handle (src/capi/capi.cpp:96-173)
template <typename T>
struct dft_obj : plan_ref
{
    std::vector<plan_desc> desc;       // the reference plan of every axis
    std::unique_ptr<gpu_dft<T>> g;     // one axis
    std::unique_ptr<gpu_dft_md<T>> md; // two or more axes
    stager st;
    explicit dft_obj(const std::vector<size_t>& shape)
    {
        const dtype_t dt = sizeof(T) == 4 ? dtype_t::f32 : dtype_t::f64;
        if (shape.empty())
            throw std::invalid_argument("a DFT plan needs at least one axis");
        for (size_t n : shape)
            desc.push_back(describe_dft(dt, n, false));
        if (shape.size() == 1)
            g.reset(new gpu_dft<T>(shape[0]));
        else
            md.reset(new gpu_dft_md<T>(shape));
    }
    size_t size() const { return g ? g->size() : md->size(); }
    size_t temp_size() const
    {
        size_t t = 0;
        for (const auto& d : desc)
            t = std::max(t, d.temp_size);
        return t;
    }
    std::string schedule() const { return g ? g->schedule().text() : md->schedule(); }
    stream_order ord;
    void execute(void* out, const void* in, long long batch, bool inverse, hipStream_t s)
    {
        ord.run(s,
                [&]
                {
                    if (g)
                    {
                        g->execute(out, in, batch, inverse, s);
                        return;
                    }
                    const size_t bytes = md->size() * 2 * sizeof(T);
                    for (long long b = 0; b < batch; ++b)
                        md->execute(static_cast<char*>(out) + b * bytes, static_cast<const char*>(in) + b * bytes,
                                    inverse, s);
                });
    }
};

// a real plan: one axis (dft_plan_real) or several (dft_plan_md_real, CCs along the last axis)
template <typename T>
struct real_obj : plan_ref
{
    std::vector<plan_desc> desc;            // complex axes first, then the real last axis
    std::unique_ptr<gpu_dft_real<T>> g;     // one axis
    std::unique_ptr<gpu_dft_md_real<T>> md; // two or more axes
    size_t extra_temp = 0;                  // md plans without real_out_is_enough keep the complex array
    stager st;
    real_obj(size_t n, int fmt)
    {
        desc.push_back(describe_dft_real(sizeof(T) == 4 ? dtype_t::f32 : dtype_t::f64, n, fmt));
        g.reset(new gpu_dft_real<T>(n, fmt));
    }
    // dft_plan_md_real(shape, real_out_is_enough), include/kfr/dft/fft.hpp:739-762
    real_obj(const std::vector<size_t>& shape, bool real_out_is_enough)
    {
        const dtype_t dt = sizeof(T) == 4 ? dtype_t::f32 : dtype_t::f64;
        if (shape.empty())
            throw std::invalid_argument("a DFT plan needs at least one axis");
        for (size_t i = 0; i + 1 < shape.size(); ++i)
            desc.push_back(describe_dft(dt, shape[i], false));
        desc.push_back(describe_dft_real(dt, shape.back(), 1));
        if (shape.size() == 1)
            g.reset(new gpu_dft_real<T>(shape[0], 1));
        else
            md.reset(new gpu_dft_md_real<T>(shape));
        if (!real_out_is_enough)
            extra_temp = complex_size() * 2 * sizeof(T);
    }
    size_t size() const { return g ? g->size() : md->size(); }
    size_t complex_size() const
    {
        if (g)
            return g->complex_size();
        return md->complex_size();
    }
    size_t temp_size() const
    {
        size_t t = 0;
        for (const auto& d : desc)
            t = std::max(t, d.temp_size);
        return t + extra_temp;
    }
    std::string schedule() const { return g ? g->schedule().text() : md->schedule(); }
    stream_order ord;
    void forward(void* out, const void* in, long long batch, hipStream_t s)
    {
        ord.run(s,
                [&]
                {
                    if (g)
                    {
                        g->forward(out, in, batch, s);
                        return;
                    }
                    for (long long b = 0; b < batch; ++b)
                        md->forward(static_cast<char*>(out) + b * complex_size() * 2 * sizeof(T),
                                    static_cast<const char*>(in) + b * size() * sizeof(T), s);
                });
    }
    void inverse(void* out, const void* in, long long batch, hipStream_t s)
    {
        ord.run(s,
                [&]
                {
                    if (g)
                    {
                        g->inverse(out, in, batch, s);
                        return;
                    }
                    for (long long b = 0; b < batch; ++b)
                        md->inverse(static_cast<char*>(out) + b * size() * sizeof(T),
                                    static_cast<const char*>(in) + b * complex_size() * 2 * sizeof(T), s);
                });
    }
};
template <typename T>
struct dct_obj : plan_ref
{
    plan_desc desc;
    gpu_dct<T> g;
    stager st;
    stream_order ord;
    explicit dct_obj(size_t n) : desc(describe_dct(sizeof(T) == 4 ? dtype_t::f32 : dtype_t::f64, n)), g(n) {}
    void execute(void* out, const void* in, long long batch, bool inverse, hipStream_t s)
    {
        ord.run(s, [&] { g.execute(out, in, batch, inverse, s); });
    }
};
template <typename T>
struct filter_obj
{
    gpu_conv_filter<T> g;
    dev_mem in_buf, out_buf;
    std::mutex mu;
    filter_obj(const T* taps, size_t n, size_t block) : g(taps, n, block) {}
};

void print_dump(const std::vector<plan_desc>& ds, const std::string& sched)
{
    for (const auto& d : ds)
        std::fputs(dump_text(d, "gfx950").c_str(), stdout);
    std::printf("%s\n", sched.c_str());
    std::fflush(stdout);
}

template <typename T>
dft_obj<T>* dft_cast(void* p)
{
    check_plan(p);
    return static_cast<dft_obj<T>*>(p);
}
template <typename T>
real_obj<T>* real_cast(void* p)
{
    check_plan(p);
    return static_cast<real_obj<T>*>(p);
}
template <typename T>
dct_obj<T>* dct_cast(void* p)
{
    check_plan(p);
    return static_cast<dct_obj<T>*>(p);
}

template <typename T>
void dft_exec(void* plan, void* out, const void* in, bool inverse)
{
    auto* o        = dft_cast<T>(plan);
    const size_t b = o->size() * 2 * sizeof(T);
    o->st.run(out, b, in, b, [&](void* dout, const void* din, hipStream_t s) { o->execute(dout, din, 1, inverse, s); });
}
template <typename T>
void real_exec(void* plan, void* out, const void* in, bool inverse)
{
    auto* o         = real_cast<T>(plan);
    const size_t rb = o->size() * sizeof(T);
    const size_t cb = o->complex_size() * 2 * sizeof(T);
    if (!inverse)
        o->st.run(out, cb, in, rb, [&](void* dout, const void* din, hipStream_t s) { o->forward(dout, din, 1, s); });
    else
        o->st.run(out, rb, in, cb, [&](void* dout, const void* din, hipStream_t s) { o->inverse(dout, din, 1, s); });
}
template <typename T>
void dct_exec(void* plan, void* out, const void* in, bool inverse)
{
    auto* o        = dct_cast<T>(plan);
    const size_t b = o->g.size() * sizeof(T);
    o->st.run(out, b, in, b, [&](void* dout, const void* din, hipStream_t s) { o->execute(dout, din, 1, inverse, s); });
}

template <typename T>
void filter_process(void* plan, T* out, const T* in, size_t n)
{
    check_plan(plan);
    auto* f = static_cast<filter_obj<T>*>(plan);
    std::lock_guard<std::mutex> lock(f->mu);
    hipStream_t s = library_stream();
    if (n == 0)
        return;
    if (is_device_ptr(in) && is_device_ptr(out) && in != out)
    {
        f->g.process(out, in, n, s);
        hip_check(hipStreamSynchronize(s), "filter process");
        return;
    }
    f->in_buf.ensure(n * sizeof(T));
    f->out_buf.ensure(n * sizeof(T));
    hip_check(hipMemcpyAsync(f->in_buf.ptr, in, n * sizeof(T), hipMemcpyDefault, s), "filter input copy");
    f->g.process(f->out_buf.template as<T>(), f->in_buf.template as<T>(), n, s);
    hip_check(hipMemcpyAsync(out, f->out_buf.ptr, n * sizeof(T), hipMemcpyDefault, s), "filter output copy");
    hip_check(hipStreamSynchronize(s), "filter process");
}

// init_shape (src/capi/capi.cpp:105-113)
std::vector<size_t> shape_of(size_t dims, const unsigned* shape)
{
    if (dims == 0 || !shape)
        throw std::invalid_argument("shape is empty");
    return std::vector<size_t>(shape, shape + dims);
}

[[noreturn]] void not_provided(const char* what)
{
    throw std::logic_error(std::string(what) + " is not provided by this library (DFT path only)");
}

// ---- process-wide plan cache (the reference's dft_cache, include/kfr/dft/cache.hpp:44-126, keys only
// on type and size); here the key also names the transform kind, the pack format, the batch layout and
// the schedule variant, so that plans of one size but different kinds never alias
struct cache_key
{
    int dtype   = 0; // 0 f32, 1 f64
    int kind    = 0; // 0 complex, 1 real, 3 real convolution, 4 complex convolution
    size_t n    = 0; // transform size (real convolution: signal length; complex convolution: FFT size)
    size_t m    = 0; // real convolution: filter length
    int fmt     = 0; // real pack format (0 Perm, 1 CCs)
    int layout  = 0; // batch layout: 0 = contiguous items
    int tuning  = 0; // schedule variant: 0 = the measured default
    auto tie() const { return std::tie(dtype, kind, n, m, fmt, layout, tuning); }
    bool operator<(const cache_key& o) const { return tie() < o.tie(); }
};

class plan_cache
{
public:
    static plan_cache& get()
    {
        static plan_cache* c = new plan_cache; // never destroyed: plans outlive static destruction
        return *c;
    }
    // a new reference to the cached object of key k, created by make() when absent
    template <typename Obj, typename Make>
    Obj* acquire(const cache_key& k, Make&& make)
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = map.find(k);
        if (it == map.end())
            it = map.emplace(k, static_cast<plan_ref*>(make())).first;
        retain(it->second);
        return static_cast<Obj*>(it->second);
    }
    void clear()
    {
        std::map<cache_key, plan_ref*> old;
        {
            std::lock_guard<std::mutex> lock(mu);
            old.swap(map);
        }
        for (auto& e : old)
            release(e.second);
    }
    size_t size()
    {
        std::lock_guard<std::mutex> lock(mu);
        return map.size();
    }

private:
    std::mutex mu;
    std::map<cache_key, plan_ref*> map;
};

// one reference held for a scope
struct ref_holder
{
    plan_ref* p;
    ~ref_holder() { release(p); }
};

// convolution plans of the cache: the real convolution of two lengths, or the complex FFT of one size,
// with their staging buffers (calls on one entry are serialised)
template <typename T>
struct conv_entry : plan_ref
{
    std::mutex mu;
    std::unique_ptr<gpu_conv<T>> real;
    std::unique_ptr<gpu_dft<T>> cplx;
    dev_mem da, db, dy, A, B;
};

template <typename T>
long convolve_impl(int mode, const void* a, size_t na, const void* b, size_t nb, void* out, bool complex_in)
{
    if (mode == 2)
    {
        b  = a;
        nb = na;
    }
    if (na == 0 || nb == 0)
        throw std::invalid_argument("convolve: empty input");
    const bool correlate = mode != 0;
    const size_t ny      = na + nb - 1;
    const size_t es      = complex_in ? 2 * sizeof(T) : sizeof(T);
    // results of autocorrelate keep the second half (include/kfr/dft/convolution.hpp:70-76)
    const size_t off  = mode == 2 ? ny / 2 : 0;
    const size_t nout = ny - off;
    if (!out)
        return (long)nout;
    // the plans come from the cache (the reference's convolve takes its plan from dft_cache,
    // src/dft/convolution-impl.cpp:106-128): repeated calls of the same sizes build no tables
    hipStream_t s  = library_stream();
    const size_t S = next_pow2(ny);
    cache_key k;
    k.dtype = sizeof(T) == 4 ? 0 : 1;
    k.kind  = complex_in ? 4 : 3;
    k.n     = complex_in ? S : na;
    k.m     = complex_in ? 0 : nb;
    auto* e = plan_cache::get().acquire<conv_entry<T>>(k,
                                                       [&]
                                                       {
                                                           auto* c = new conv_entry<T>();
                                                           try
                                                           {
                                                               if (complex_in)
                                                                   c->cplx.reset(new gpu_dft<T>(S));
                                                               else
                                                                   c->real.reset(new gpu_conv<T>(na, nb));
                                                           }
                                                           catch (...)
                                                           {
                                                               delete c;
                                                               throw;
                                                           }
                                                           return c;
                                                       });
    ref_holder hold{ e };
    std::lock_guard<std::mutex> lock(e->mu);
    e->da.ensure(na * es);
    e->db.ensure(nb * es);
    e->dy.ensure(ny * es);
    hip_check(hipMemcpyAsync(e->da.ptr, a, na * es, hipMemcpyDefault, s), "convolve input copy");
    hip_check(hipMemcpyAsync(e->db.ptr, b, nb * es, hipMemcpyDefault, s), "convolve input copy");
    if (complex_in)
    {
        e->A.ensure(S * es);
        e->B.ensure(S * es);
        gpu_convolve_complex<T>(e->dy.ptr, e->da.ptr, na, e->db.ptr, nb, correlate, false, *e->cplx, e->A.ptr, e->B.ptr,
                                s);
    }
    else
        e->real->forward(e->dy.ptr, e->da.ptr, e->db.ptr, 0, 1, correlate, s);
    hip_check(hipMemcpyAsync(out, static_cast<char*>(e->dy.ptr) + off * es, nout * es, hipMemcpyDefault, s),
              "convolve output copy");
    hip_check(hipStreamSynchronize(s), "convolve");
    return (long)nout;
}

// streaming filter of complex signals (convolve_filter<complex<T>>)
template <typename T>
struct cfilter_obj
{
    gpu_conv_filter_complex<T> g;
    dev_mem in_buf, out_buf;
    std::mutex mu;
    cfilter_obj(const void* taps, size_t n, size_t block) : g(taps, n, block) {}
};

template <typename T>
void cfilter_process(void* plan, void* out, const void* in, size_t n)
{
    check_plan(plan);
    auto* f = static_cast<cfilter_obj<T>*>(plan);
    std::lock_guard<std::mutex> lock(f->mu);
    if (n == 0)
        return;
    hipStream_t s      = library_stream();
    const size_t bytes = n * 2 * sizeof(T);
    f->in_buf.ensure(bytes);
    f->out_buf.ensure(bytes);
    hip_check(hipMemcpyAsync(f->in_buf.ptr, in, bytes, hipMemcpyDefault, s), "filter input copy");
    f->g.process(f->out_buf.ptr, f->in_buf.ptr, n, s);
    hip_check(hipMemcpyAsync(out, f->out_buf.ptr, bytes, hipMemcpyDefault, s), "filter output copy");
    hip_check(hipStreamSynchronize(s), "filter process");
}

// ---- progressive execution (include/kfr/dft/fft.hpp:393-439): progressive_total_steps() steps as the
// reference counts them (its stage count for this plan and direction); the GPU units of the plan are
// spread over the steps (unit u at step round(u (K-1) / (U-1)), a single unit at step 0), the input is
// read at the first step and the output written at the last. Each step is synchronous.
struct progressive_obj
{
    virtual ~progressive_obj() = default;
    size_t total = 1, step = 0;
    int units    = 1;
    int unit_step(int u) const
    {
        if (units <= 1 || total <= 1)
            return 0;
        return (int)(((long long)u * (long long)(total - 1) * 2 + (units - 1)) / (2 * (long long)(units - 1)));
    }
    virtual void run_step(size_t i) = 0;
};

template <typename T>
struct dft_progressive : progressive_obj
{
    dft_obj<T>* plan;
    bool inverse;
    void* out;
    const void* in;
    dev_mem W, mid;
    dft_progressive(dft_obj<T>* p, bool prog_opt, bool inv, void* o, const void* i)
        : plan(p), inverse(inv), out(o), in(i)
    {
        if (!plan->g)
            throw std::logic_error("progressive execution is provided for one-dimensional plans");
        const dtype_t dt = sizeof(T) == 4 ? dtype_t::f32 : dtype_t::f64;
        total            = std::max<size_t>(1, describe_dft(dt, plan->size(), prog_opt).order[inverse ? 1 : 0].size());
        units            = plan->g->progressive_prepare((int)total);
        W.reset(plan->size() * 2 * sizeof(T));
        if (size_t mb = plan->g->progressive_mid_bytes())
            mid.reset(mb);
        retain(plan);
    }
    ~dft_progressive() override { release(plan); }
    void run_step(size_t i) override
    {
        hipStream_t s      = library_stream();
        const size_t bytes = plan->size() * 2 * sizeof(T);
        if (i == 0)
            hip_check(hipDeviceSynchronize(), "progressive start"); // device inputs may still be in production
        plan->ord.run(s,
                      [&]
                      {
                          if (i == 0)
                              hip_check(hipMemcpyAsync(W.ptr, in, bytes, hipMemcpyDefault, s), "progressive input copy");
                          for (int u = 0; u < units; ++u)
                              if ((size_t)unit_step(u) == i)
                                  plan->g->progressive_unit(u, W.ptr, mid.ptr, inverse, s);
                          if (i + 1 == total)
                              hip_check(hipMemcpyAsync(out, W.ptr, bytes, hipMemcpyDefault, s), "progressive output copy");
                      });
        hip_check(hipStreamSynchronize(s), "progressive step");
    }
};

template <typename T>
struct real_progressive : progressive_obj
{
    real_obj<T>* plan;
    bool inverse;
    void* out;
    const void* in;
    dev_mem Z, Cb, mid;
    real_progressive(real_obj<T>* p, bool inv, void* o, const void* i) : plan(p), inverse(inv), out(o), in(i)
    {
        if (!plan->g)
            throw std::logic_error("progressive execution is provided for one-dimensional plans");
        const dtype_t dt = sizeof(T) == 4 ? dtype_t::f32 : dtype_t::f64;
        total = std::max<size_t>(1, describe_dft_real(dt, plan->size(), plan->g->format()).order[inverse ? 1 : 0].size());
        units = plan->g->progressive_prepare((int)total);
        Z.reset(plan->size() * sizeof(T));
        Cb.reset(plan->complex_size() * 2 * sizeof(T));
        if (size_t mb = plan->g->progressive_mid_bytes())
            mid.reset(mb);
        retain(plan);
    }
    ~real_progressive() override { release(plan); }
    void run_step(size_t i) override
    {
        hipStream_t s     = library_stream();
        const size_t rb   = plan->size() * sizeof(T);
        const size_t cb   = plan->complex_size() * 2 * sizeof(T);
        if (i == 0)
            hip_check(hipDeviceSynchronize(), "progressive start");
        plan->ord.run(s,
                      [&]
                      {
                          if (i == 0)
                              hip_check(hipMemcpyAsync(inverse ? Cb.ptr : Z.ptr, in, inverse ? cb : rb, hipMemcpyDefault,
                                                       s),
                                        "progressive input copy");
                          for (int u = 0; u < units; ++u)
                              if ((size_t)unit_step(u) == i)
                                  plan->g->progressive_unit(u, Z.ptr, Cb.ptr, mid.ptr, inverse, s);
                          if (i + 1 == total)
                              hip_check(hipMemcpyAsync(out, inverse ? Z.ptr : Cb.ptr, inverse ? rb : cb, hipMemcpyDefault,
                                                       s),
                                        "progressive output copy");
                      });
        hip_check(hipStreamSynchronize(s), "progressive step");
    }
};

struct conv_obj
{
    int dtype;
    std::unique_ptr<gpu_conv<float>> f;
    std::unique_ptr<gpu_conv<double>> d;
    stream_order ord;
};

struct batch_obj
{
    std::unique_ptr<gpu_batch<float>> f;
    std::unique_ptr<gpu_batch<double>> d;
    dev_mem buf; // staging for host pointers
    std::mutex mu;
    stream_order ord;
    size_t total() const { return f ? f->total() : d->total(); }
    size_t elem() const { return f ? sizeof(float) * 2 : sizeof(double) * 2; }
    void run(void* out, const void* in, hipStream_t s)
    {
        ord.run(s,
                [&]
                {
                    if (f)
                        f->execute(out, in, s);
                    else
                        d->execute(out, in, s);
                });
    }
};

struct half_obj
{
    std::unique_ptr<pow2_fft<float>> f;
    size_t n = 0;
    dev_mem buf; // staging for host pointers
    std::mutex mu;
    stream_order ord;
    void run(bool inverse, const void* in, void* out, long long batch, hipStream_t s)
    {
        ord.run(s, [&] { f->run(inverse, in, out, batch, s); });
    }
};

int copy_text(const std::string& s, char* buf, size_t buflen)
{
    if (buf && buflen)
    {
        const size_t n = std::min(buflen - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return (int)s.size();
}

dtype_t to_dtype(int d)
{
    if (d != 0 && d != 1)
        throw std::invalid_argument("dtype must be 0 (f32) or 1 (f64)");
    return d == 0 ? dtype_t::f32 : dtype_t::f64;
}

} // namespace

extern "C"
{

    const char* kfr_version_string(void) { return "KFR 7.0.1 gfx950 64-bit (hip/MI355X) kfr_amd"; }
    uint32_t kfr_version(void) { return 7 * 10000 + 0 * 100 + 1; }
    const char* kfr_enabled_archs(void) { return "gfx950"; }
    int kfr_current_arch(void) { return KFR_ARCH_GFX950; }
    const char* kfr_last_error(void) { return latest_error.data(); }

    void* kfr_allocate(size_t size) { return kfr_allocate_aligned(size, KFR_DEFAULT_ALIGNMENT); }
    void* kfr_allocate_aligned(size_t size, size_t alignment)
    {
        // one header slot before the block keeps the base pointer for kfr_deallocate
        if (alignment < sizeof(void*) || (alignment & (alignment - 1)))
            alignment = KFR_DEFAULT_ALIGNMENT;
        void* raw = std::malloc(size + alignment + sizeof(void*));
        if (!raw)
            return nullptr;
        uintptr_t p = (reinterpret_cast<uintptr_t>(raw) + sizeof(void*) + alignment - 1) & ~(uintptr_t)(alignment - 1);
        reinterpret_cast<void**>(p)[-1] = raw;
        return reinterpret_cast<void*>(p);
    }
    void kfr_deallocate(void* ptr)
    {
        if (ptr)
            std::free(reinterpret_cast<void**>(ptr)[-1]);
    }

    // ---- complex DFT
    KFR_DFT_PLAN_F32* kfr_dft_create_plan_f32(size_t size)
    {
        return try_fn([&] { return reinterpret_cast<KFR_DFT_PLAN_F32*>(new dft_obj<float>(std::vector<size_t>{ size })); },
                      (KFR_DFT_PLAN_F32*)nullptr);
    }
    KFR_DFT_PLAN_F64* kfr_dft_create_plan_f64(size_t size)
    {
        return try_fn([&] { return reinterpret_cast<KFR_DFT_PLAN_F64*>(new dft_obj<double>(std::vector<size_t>{ size })); },
                      (KFR_DFT_PLAN_F64*)nullptr);
    }
    KFR_DFT_PLAN_F32* kfr_dft_create_2d_plan_f32(size_t size1, size_t size2)
    {
        return try_fn([&] { return reinterpret_cast<KFR_DFT_PLAN_F32*>(new dft_obj<float>(std::vector<size_t>{ size1, size2 })); },
                      (KFR_DFT_PLAN_F32*)nullptr);
    }
    KFR_DFT_PLAN_F64* kfr_dft_create_2d_plan_f64(size_t size1, size_t size2)
    {
        return try_fn([&] { return reinterpret_cast<KFR_DFT_PLAN_F64*>(new dft_obj<double>(std::vector<size_t>{ size1, size2 })); },
                      (KFR_DFT_PLAN_F64*)nullptr);
    }
    KFR_DFT_PLAN_F32* kfr_dft_create_3d_plan_f32(size_t size1, size_t size2, size_t size3)
    {
        return try_fn(
            [&] { return reinterpret_cast<KFR_DFT_PLAN_F32*>(new dft_obj<float>(std::vector<size_t>{ size1, size2, size3 })); },
            (KFR_DFT_PLAN_F32*)nullptr);
    }
    KFR_DFT_PLAN_F64* kfr_dft_create_3d_plan_f64(size_t size1, size_t size2, size_t size3)
    {
        return try_fn(
            [&] { return reinterpret_cast<KFR_DFT_PLAN_F64*>(new dft_obj<double>(std::vector<size_t>{ size1, size2, size3 })); },
            (KFR_DFT_PLAN_F64*)nullptr);
    }
    KFR_DFT_PLAN_F32* kfr_dft_create_md_plan_f32(size_t dims, const unsigned* shape)
    {
        return try_fn([&] { return reinterpret_cast<KFR_DFT_PLAN_F32*>(new dft_obj<float>(shape_of(dims, shape))); },
                      (KFR_DFT_PLAN_F32*)nullptr);
    }
    KFR_DFT_PLAN_F64* kfr_dft_create_md_plan_f64(size_t dims, const unsigned* shape)
    {
        return try_fn([&] { return reinterpret_cast<KFR_DFT_PLAN_F64*>(new dft_obj<double>(shape_of(dims, shape))); },
                      (KFR_DFT_PLAN_F64*)nullptr);
    }
    void kfr_dft_dump_f32(KFR_DFT_PLAN_F32* plan)
    {
        try_void([&] { print_dump(dft_cast<float>(plan)->desc, dft_cast<float>(plan)->schedule()); });
    }
    void kfr_dft_dump_f64(KFR_DFT_PLAN_F64* plan)
    {
        try_void([&] { print_dump(dft_cast<double>(plan)->desc, dft_cast<double>(plan)->schedule()); });
    }
    size_t kfr_dft_get_size_f32(KFR_DFT_PLAN_F32* plan)
    {
        return try_fn([&] { return dft_cast<float>(plan)->size(); }, (size_t)0);
    }
    size_t kfr_dft_get_size_f64(KFR_DFT_PLAN_F64* plan)
    {
        return try_fn([&] { return dft_cast<double>(plan)->size(); }, (size_t)0);
    }
    size_t kfr_dft_get_temp_size_f32(KFR_DFT_PLAN_F32* plan)
    {
        return try_fn([&] { return dft_cast<float>(plan)->temp_size(); }, (size_t)0);
    }
    size_t kfr_dft_get_temp_size_f64(KFR_DFT_PLAN_F64* plan)
    {
        return try_fn([&] { return dft_cast<double>(plan)->temp_size(); }, (size_t)0);
    }
    void kfr_dft_execute_f32(KFR_DFT_PLAN_F32* plan, kfr_c32* out, const kfr_c32* in, uint8_t*)
    {
        try_void([&] { dft_exec<float>(plan, out, in, false); });
    }
    void kfr_dft_execute_f64(KFR_DFT_PLAN_F64* plan, kfr_c64* out, const kfr_c64* in, uint8_t*)
    {
        try_void([&] { dft_exec<double>(plan, out, in, false); });
    }
    void kfr_dft_execute_inverse_f32(KFR_DFT_PLAN_F32* plan, kfr_c32* out, const kfr_c32* in, uint8_t*)
    {
        try_void([&] { dft_exec<float>(plan, out, in, true); });
    }
    void kfr_dft_execute_inverse_f64(KFR_DFT_PLAN_F64* plan, kfr_c64* out, const kfr_c64* in, uint8_t*)
    {
        try_void([&] { dft_exec<double>(plan, out, in, true); });
    }
    void kfr_dft_delete_plan_f32(KFR_DFT_PLAN_F32* plan)
    {
        try_void([&] { release(reinterpret_cast<dft_obj<float>*>(plan)); });
    }
    void kfr_dft_delete_plan_f64(KFR_DFT_PLAN_F64* plan)
    {
        try_void([&] { release(reinterpret_cast<dft_obj<double>*>(plan)); });
    }

    // ---- real DFT
    KFR_DFT_REAL_PLAN_F32* kfr_dft_real_create_plan_f32(size_t size, KFR_DFT_PACK_FORMAT fmt)
    {
        return try_fn([&] { return reinterpret_cast<KFR_DFT_REAL_PLAN_F32*>(new real_obj<float>(size, (int)fmt)); },
                      (KFR_DFT_REAL_PLAN_F32*)nullptr);
    }
    KFR_DFT_REAL_PLAN_F64* kfr_dft_real_create_plan_f64(size_t size, KFR_DFT_PACK_FORMAT fmt)
    {
        return try_fn([&] { return reinterpret_cast<KFR_DFT_REAL_PLAN_F64*>(new real_obj<double>(size, (int)fmt)); },
                      (KFR_DFT_REAL_PLAN_F64*)nullptr);
    }
    KFR_DFT_REAL_PLAN_F32* kfr_dft_real_create_2d_plan_f32(size_t size1, size_t size2, kfr_bool real_out_is_enough)
    {
        return try_fn(
            [&] {
                return reinterpret_cast<KFR_DFT_REAL_PLAN_F32*>(
                    new real_obj<float>(std::vector<size_t>{ size1, size2 }, (bool)real_out_is_enough));
            },
            (KFR_DFT_REAL_PLAN_F32*)nullptr);
    }
    KFR_DFT_REAL_PLAN_F64* kfr_dft_real_create_2d_plan_f64(size_t size1, size_t size2, kfr_bool real_out_is_enough)
    {
        return try_fn(
            [&] {
                return reinterpret_cast<KFR_DFT_REAL_PLAN_F64*>(
                    new real_obj<double>(std::vector<size_t>{ size1, size2 }, (bool)real_out_is_enough));
            },
            (KFR_DFT_REAL_PLAN_F64*)nullptr);
    }
    KFR_DFT_REAL_PLAN_F32* kfr_dft_real_create_3d_plan_f32(size_t size1, size_t size2, size_t size3, kfr_bool real_out_is_enough)
    {
        return try_fn(
            [&] {
                return reinterpret_cast<KFR_DFT_REAL_PLAN_F32*>(
                    new real_obj<float>(std::vector<size_t>{ size1, size2, size3 }, (bool)real_out_is_enough));
            },
            (KFR_DFT_REAL_PLAN_F32*)nullptr);
    }
    KFR_DFT_REAL_PLAN_F64* kfr_dft_real_create_3d_plan_f64(size_t size1, size_t size2, size_t size3, kfr_bool real_out_is_enough)
    {
        return try_fn(
            [&] {
                return reinterpret_cast<KFR_DFT_REAL_PLAN_F64*>(
                    new real_obj<double>(std::vector<size_t>{ size1, size2, size3 }, (bool)real_out_is_enough));
            },
            (KFR_DFT_REAL_PLAN_F64*)nullptr);
    }
    KFR_DFT_REAL_PLAN_F32* kfr_dft_real_create_md_plan_f32(size_t dims, const unsigned* shape, kfr_bool real_out_is_enough)
    {
        return try_fn(
            [&] { return reinterpret_cast<KFR_DFT_REAL_PLAN_F32*>(new real_obj<float>(shape_of(dims, shape), (bool)real_out_is_enough)); },
            (KFR_DFT_REAL_PLAN_F32*)nullptr);
    }
    KFR_DFT_REAL_PLAN_F64* kfr_dft_real_create_md_plan_f64(size_t dims, const unsigned* shape, kfr_bool real_out_is_enough)
    {
        return try_fn(
            [&] { return reinterpret_cast<KFR_DFT_REAL_PLAN_F64*>(new real_obj<double>(shape_of(dims, shape), (bool)real_out_is_enough)); },
            (KFR_DFT_REAL_PLAN_F64*)nullptr);
    }
    void kfr_dft_real_dump_f32(KFR_DFT_REAL_PLAN_F32* plan)
    {
        try_void([&] { print_dump(real_cast<float>(plan)->desc, real_cast<float>(plan)->schedule()); });
    }
    void kfr_dft_real_dump_f64(KFR_DFT_REAL_PLAN_F64* plan)
    {
        try_void([&] { print_dump(real_cast<double>(plan)->desc, real_cast<double>(plan)->schedule()); });
    }
    size_t kfr_dft_real_get_size_f32(KFR_DFT_REAL_PLAN_F32* plan)
    {
        return try_fn([&] { return real_cast<float>(plan)->size(); }, (size_t)0);
    }
    size_t kfr_dft_real_get_size_f64(KFR_DFT_REAL_PLAN_F64* plan)
    {
        return try_fn([&] { return real_cast<double>(plan)->size(); }, (size_t)0);
    }
    size_t kfr_dft_real_get_temp_size_f32(KFR_DFT_REAL_PLAN_F32* plan)
    {
        return try_fn([&] { return real_cast<float>(plan)->temp_size(); }, (size_t)0);
    }
    size_t kfr_dft_real_get_temp_size_f64(KFR_DFT_REAL_PLAN_F64* plan)
    {
        return try_fn([&] { return real_cast<double>(plan)->temp_size(); }, (size_t)0);
    }
    void kfr_dft_real_execute_f32(KFR_DFT_REAL_PLAN_F32* plan, kfr_c32* out, const kfr_f32* in, uint8_t*)
    {
        try_void([&] { real_exec<float>(plan, out, in, false); });
    }
    void kfr_dft_real_execute_f64(KFR_DFT_REAL_PLAN_F64* plan, kfr_c64* out, const kfr_f64* in, uint8_t*)
    {
        try_void([&] { real_exec<double>(plan, out, in, false); });
    }
    void kfr_dft_real_execute_inverse_f32(KFR_DFT_REAL_PLAN_F32* plan, kfr_f32* out, const kfr_c32* in, uint8_t*)
    {
        try_void([&] { real_exec<float>(plan, out, in, true); });
    }
    void kfr_dft_real_execute_inverse_f64(KFR_DFT_REAL_PLAN_F64* plan, kfr_f64* out, const kfr_c64* in, uint8_t*)
    {
        try_void([&] { real_exec<double>(plan, out, in, true); });
    }
    void kfr_dft_real_delete_plan_f32(KFR_DFT_REAL_PLAN_F32* plan)
    {
        try_void([&] { release(reinterpret_cast<real_obj<float>*>(plan)); });
    }
    void kfr_dft_real_delete_plan_f64(KFR_DFT_REAL_PLAN_F64* plan)
    {
        try_void([&] { release(reinterpret_cast<real_obj<double>*>(plan)); });
    }

    // ---- DCT
    KFR_DCT_PLAN_F32* kfr_dct_create_plan_f32(size_t size)
    {
        return try_fn([&] { return reinterpret_cast<KFR_DCT_PLAN_F32*>(new dct_obj<float>(size)); },
                      (KFR_DCT_PLAN_F32*)nullptr);
    }
    KFR_DCT_PLAN_F64* kfr_dct_create_plan_f64(size_t size)
    {
        return try_fn([&] { return reinterpret_cast<KFR_DCT_PLAN_F64*>(new dct_obj<double>(size)); },
                      (KFR_DCT_PLAN_F64*)nullptr);
    }
    void kfr_dct_dump_f32(KFR_DCT_PLAN_F32* plan)
    {
        try_void([&] { print_dump({ dct_cast<float>(plan)->desc }, dct_cast<float>(plan)->g.schedule().text()); });
    }
    void kfr_dct_dump_f64(KFR_DCT_PLAN_F64* plan)
    {
        try_void([&] { print_dump({ dct_cast<double>(plan)->desc }, dct_cast<double>(plan)->g.schedule().text()); });
    }
    size_t kfr_dct_get_size_f32(KFR_DCT_PLAN_F32* plan)
    {
        return try_fn([&] { return dct_cast<float>(plan)->g.size(); }, (size_t)0);
    }
    size_t kfr_dct_get_size_f64(KFR_DCT_PLAN_F64* plan)
    {
        return try_fn([&] { return dct_cast<double>(plan)->g.size(); }, (size_t)0);
    }
    size_t kfr_dct_get_temp_size_f32(KFR_DCT_PLAN_F32* plan)
    {
        return try_fn([&] { return dct_cast<float>(plan)->desc.temp_size; }, (size_t)0);
    }
    size_t kfr_dct_get_temp_size_f64(KFR_DCT_PLAN_F64* plan)
    {
        return try_fn([&] { return dct_cast<double>(plan)->desc.temp_size; }, (size_t)0);
    }
    void kfr_dct_execute_f32(KFR_DCT_PLAN_F32* plan, kfr_f32* out, const kfr_f32* in, uint8_t*)
    {
        try_void([&] { dct_exec<float>(plan, out, in, false); });
    }
    void kfr_dct_execute_f64(KFR_DCT_PLAN_F64* plan, kfr_f64* out, const kfr_f64* in, uint8_t*)
    {
        try_void([&] { dct_exec<double>(plan, out, in, false); });
    }
    void kfr_dct_execute_inverse_f32(KFR_DCT_PLAN_F32* plan, kfr_f32* out, const kfr_f32* in, uint8_t*)
    {
        try_void([&] { dct_exec<float>(plan, out, in, true); });
    }
    void kfr_dct_execute_inverse_f64(KFR_DCT_PLAN_F64* plan, kfr_f64* out, const kfr_f64* in, uint8_t*)
    {
        try_void([&] { dct_exec<double>(plan, out, in, true); });
    }
    void kfr_dct_delete_plan_f32(KFR_DCT_PLAN_F32* plan)
    {
        try_void([&] { release(reinterpret_cast<dct_obj<float>*>(plan)); });
    }
    void kfr_dct_delete_plan_f64(KFR_DCT_PLAN_F64* plan)
    {
        try_void([&] { release(reinterpret_cast<dct_obj<double>*>(plan)); });
    }

    // ---- filters
    KFR_FILTER_F32* kfr_filter_create_fir_plan_f32(const kfr_f32*, size_t)
    {
        return try_fn([&]() -> KFR_FILTER_F32* { not_provided("FIR filter"); }, (KFR_FILTER_F32*)nullptr);
    }
    KFR_FILTER_F64* kfr_filter_create_fir_plan_f64(const kfr_f64*, size_t)
    {
        return try_fn([&]() -> KFR_FILTER_F64* { not_provided("FIR filter"); }, (KFR_FILTER_F64*)nullptr);
    }
    KFR_FILTER_F32* kfr_filter_create_iir_plan_f32(const kfr_f32*, size_t)
    {
        return try_fn([&]() -> KFR_FILTER_F32* { not_provided("IIR filter"); }, (KFR_FILTER_F32*)nullptr);
    }
    KFR_FILTER_F64* kfr_filter_create_iir_plan_f64(const kfr_f64*, size_t)
    {
        return try_fn([&]() -> KFR_FILTER_F64* { not_provided("IIR filter"); }, (KFR_FILTER_F64*)nullptr);
    }
    KFR_FILTER_F32* kfr_filter_create_convolution_plan_f32(const kfr_f32* taps, size_t size, size_t block_size)
    {
        return try_fn(
            [&] {
                return reinterpret_cast<KFR_FILTER_F32*>(
                    new filter_obj<float>(taps, size, block_size ? block_size : 1024));
            },
            (KFR_FILTER_F32*)nullptr);
    }
    KFR_FILTER_F64* kfr_filter_create_convolution_plan_f64(const kfr_f64* taps, size_t size, size_t block_size)
    {
        return try_fn(
            [&] {
                return reinterpret_cast<KFR_FILTER_F64*>(
                    new filter_obj<double>(taps, size, block_size ? block_size : 1024));
            },
            (KFR_FILTER_F64*)nullptr);
    }
    void kfr_filter_process_f32(KFR_FILTER_F32* plan, kfr_f32* output, const kfr_f32* input, size_t size)
    {
        try_void([&] { filter_process<float>(plan, output, input, size); });
    }
    void kfr_filter_process_f64(KFR_FILTER_F64* plan, kfr_f64* output, const kfr_f64* input, size_t size)
    {
        try_void([&] { filter_process<double>(plan, output, input, size); });
    }
    void kfr_filter_reset_f32(KFR_FILTER_F32* plan)
    {
        try_void(
            [&]
            {
                check_plan(plan);
                reinterpret_cast<filter_obj<float>*>(plan)->g.reset();
            });
    }
    void kfr_filter_reset_f64(KFR_FILTER_F64* plan)
    {
        try_void(
            [&]
            {
                check_plan(plan);
                reinterpret_cast<filter_obj<double>*>(plan)->g.reset();
            });
    }
    void kfr_filter_delete_plan_f32(KFR_FILTER_F32* plan)
    {
        try_void([&] { delete reinterpret_cast<filter_obj<float>*>(plan); });
    }
    void kfr_filter_delete_plan_f64(KFR_FILTER_F64* plan)
    {
        try_void([&] { delete reinterpret_cast<filter_obj<double>*>(plan); });
    }

    // ---- extensions
    int kfr_amd_dft_execute_batch_f32(KFR_DFT_PLAN_F32* plan, kfr_c32* out, const kfr_c32* in, size_t batch,
                                      int inverse, void* stream)
    {
        return try_fn(
            [&]
            {
                dft_cast<float>(plan)->execute(out, in, (long long)batch, inverse != 0, as_stream(stream));
                return 0;
            },
            -1);
    }
    int kfr_amd_dft_execute_batch_f64(KFR_DFT_PLAN_F64* plan, kfr_c64* out, const kfr_c64* in, size_t batch,
                                      int inverse, void* stream)
    {
        return try_fn(
            [&]
            {
                dft_cast<double>(plan)->execute(out, in, (long long)batch, inverse != 0, as_stream(stream));
                return 0;
            },
            -1);
    }
    int kfr_amd_dft_real_execute_batch_f32(KFR_DFT_REAL_PLAN_F32* plan, void* out, const void* in, size_t batch,
                                           int inverse, void* stream)
    {
        return try_fn(
            [&]
            {
                auto* o = real_cast<float>(plan);
                if (inverse)
                    o->inverse(out, in, (long long)batch, as_stream(stream));
                else
                    o->forward(out, in, (long long)batch, as_stream(stream));
                return 0;
            },
            -1);
    }
    int kfr_amd_dft_real_execute_batch_f64(KFR_DFT_REAL_PLAN_F64* plan, void* out, const void* in, size_t batch,
                                           int inverse, void* stream)
    {
        return try_fn(
            [&]
            {
                auto* o = real_cast<double>(plan);
                if (inverse)
                    o->inverse(out, in, (long long)batch, as_stream(stream));
                else
                    o->forward(out, in, (long long)batch, as_stream(stream));
                return 0;
            },
            -1);
    }
    int kfr_amd_dct_execute_batch_f32(KFR_DCT_PLAN_F32* plan, kfr_f32* out, const kfr_f32* in, size_t batch,
                                      int inverse, void* stream)
    {
        return try_fn(
            [&]
            {
                dct_cast<float>(plan)->execute(out, in, (long long)batch, inverse != 0, as_stream(stream));
                return 0;
            },
            -1);
    }
    int kfr_amd_dct_execute_batch_f64(KFR_DCT_PLAN_F64* plan, kfr_f64* out, const kfr_f64* in, size_t batch,
                                      int inverse, void* stream)
    {
        return try_fn(
            [&]
            {
                dct_cast<double>(plan)->execute(out, in, (long long)batch, inverse != 0, as_stream(stream));
                return 0;
            },
            -1);
    }
    int kfr_amd_dft_execute_batch_noalloc_f32(KFR_DFT_PLAN_F32* plan, kfr_c32* out, const kfr_c32* in, size_t batch,
                                              int inverse, void* stream)
    {
        return try_fn(
            [&]
            {
                auto* o = dft_cast<float>(plan);
                if (!o->g)
                    throw std::invalid_argument("the allocation-free execute takes one-dimensional plans");
                o->g->execute_workspace(out, in, (long long)batch, inverse != 0, as_stream(stream),
                                        noalloc_workspace());
                return 0;
            },
            -1);
    }
    size_t kfr_amd_dft_real_get_complex_size_f32(KFR_DFT_REAL_PLAN_F32* plan)
    {
        return try_fn([&] { return real_cast<float>(plan)->complex_size(); }, (size_t)0);
    }
    size_t kfr_amd_dft_real_get_complex_size_f64(KFR_DFT_REAL_PLAN_F64* plan)
    {
        return try_fn([&] { return real_cast<double>(plan)->complex_size(); }, (size_t)0);
    }

    long kfr_amd_convolve(int dtype, int mode, const void* a, size_t na, const void* b, size_t nb, void* out)
    {
        return try_fn(
            [&]() -> long
            {
                if (mode < 0 || mode > 2)
                    throw std::invalid_argument("mode must be 0 (convolve), 1 (correlate) or 2 (autocorrelate)");
                switch (dtype)
                {
                case 0:
                    return convolve_impl<float>(mode, a, na, b, nb, out, false);
                case 1:
                    return convolve_impl<double>(mode, a, na, b, nb, out, false);
                case 2:
                    return convolve_impl<float>(mode, a, na, b, nb, out, true);
                case 3:
                    return convolve_impl<double>(mode, a, na, b, nb, out, true);
                default:
                    throw std::invalid_argument("dtype must be 0..3");
                }
            },
            -1L);
    }

    KFR_AMD_CONV* kfr_amd_conv_create(int dtype, size_t nx, size_t nh)
    {
        return try_fn(
            [&]
            {
                auto* c  = new conv_obj();
                c->dtype = (int)to_dtype(dtype);
                try
                {
                    if (dtype == 0)
                        c->f.reset(new gpu_conv<float>(nx, nh));
                    else
                        c->d.reset(new gpu_conv<double>(nx, nh));
                }
                catch (...)
                {
                    delete c;
                    throw;
                }
                return reinterpret_cast<KFR_AMD_CONV*>(c);
            },
            (KFR_AMD_CONV*)nullptr);
    }
    size_t kfr_amd_conv_fft_size(KFR_AMD_CONV* conv)
    {
        return try_fn(
            [&]
            {
                check_plan(conv);
                auto* c = reinterpret_cast<conv_obj*>(conv);
                return c->f ? c->f->fft_size() : c->d->fft_size();
            },
            (size_t)0);
    }
    int kfr_amd_conv_forward(KFR_AMD_CONV* conv, void* y, const void* x, const void* h, size_t h_dist, size_t batch,
                             int correlate, void* stream)
    {
        return try_fn(
            [&]
            {
                check_plan(conv);
                auto* c       = reinterpret_cast<conv_obj*>(conv);
                hipStream_t s = as_stream(stream);
                c->ord.run(s,
                           [&]
                           {
                               if (c->f)
                                   c->f->forward(y, x, h, (long long)h_dist, (long long)batch, correlate != 0, s);
                               else
                                   c->d->forward(y, x, h, (long long)h_dist, (long long)batch, correlate != 0, s);
                           });
                return 0;
            },
            -1);
    }
    int kfr_amd_conv_backward(KFR_AMD_CONV* conv, void* dx, void* dh, const void* g, const void* x, const void* h,
                              size_t h_dist, size_t batch, int correlate, void* stream)
    {
        return try_fn(
            [&]
            {
                check_plan(conv);
                auto* c       = reinterpret_cast<conv_obj*>(conv);
                hipStream_t s = as_stream(stream);
                c->ord.run(s,
                           [&]
                           {
                               if (c->f)
                                   c->f->backward(dx, dh, g, x, h, (long long)h_dist, (long long)batch,
                                                  correlate != 0, s);
                               else
                                   c->d->backward(dx, dh, g, x, h, (long long)h_dist, (long long)batch,
                                                  correlate != 0, s);
                           });
                return 0;
            },
            -1);
    }
    void kfr_amd_conv_delete(KFR_AMD_CONV* conv)
    {
        try_void([&] { delete reinterpret_cast<conv_obj*>(conv); });
    }

    int kfr_amd_describe(int kind, int dtype, size_t size, int arg, char* buf, size_t buflen)
    {
        return try_fn(
            [&]
            {
                const dtype_t t = to_dtype(dtype);
                plan_desc d;
                switch (kind)
                {
                case 0:
                    d = describe_dft(t, size, false);
                    break;
                case 1:
                    d = describe_dft_real(t, size, arg);
                    break;
                case 2:
                    d = describe_dct(t, size);
                    break;
                case 3:
                    d = describe_dft(t, size, true);
                    break;
                default:
                    throw std::invalid_argument("kind must be 0..3");
                }
                char line[256];
                std::snprintf(line, sizeof(line), "size=%zu temp_size=%zu data_size=%zu stages=%zu,%zu arblen=%d\n",
                              d.size, d.temp_size, d.data_size, d.order[0].size(), d.order[1].size(), (int)d.arblen);
                std::string s = line;
                s += dump_text(d, "gfx950");
                for (int dir = 0; dir < 2; ++dir)
                {
                    s += dir == 0 ? "fwd:" : "inv:";
                    for (int i : d.order[dir])
                        s += " " + std::to_string(i);
                    s += "\n";
                }
                return copy_text(s, buf, buflen);
            },
            -1);
    }

    int kfr_amd_dft_permutation(int dtype, size_t size, int64_t* perm)
    {
        return try_fn(
            [&]
            {
                const auto p = reorder_permutation(describe_dft(to_dtype(dtype), size, false));
                if (p.empty())
                    return 0;
                std::copy(p.begin(), p.end(), perm);
                return 1;
            },
            -1);
    }

    size_t kfr_amd_dft_progressive_total_steps(int dtype, size_t size)
    {
        return try_fn([&] { return describe_dft(to_dtype(dtype), size, false).order[0].size(); }, (size_t)0);
    }

    int kfr_amd_dft_schedule(int dtype, size_t size, char* buf, size_t buflen)
    {
        return try_fn(
            [&]
            {
                if (to_dtype(dtype) == dtype_t::f32)
                    return copy_text(gpu_dft<float>(size).schedule().text(), buf, buflen);
                return copy_text(gpu_dft<double>(size).schedule().text(), buf, buflen);
            },
            -1);
    }

    KFR_AMD_BATCH* kfr_amd_batch_create(int dtype, size_t count, const size_t* sizes, const int* inverse,
                                        int autotune)
    {
        return try_fn(
            [&]
            {
                const dtype_t dt = to_dtype(dtype);
                if (count && !sizes)
                    throw std::invalid_argument("sizes is NULL");
                auto* b = new batch_obj();
                try
                {
                    if (dt == dtype_t::f32)
                        b->f.reset(new gpu_batch<float>(count, sizes, inverse, autotune != 0));
                    else
                        b->d.reset(new gpu_batch<double>(count, sizes, inverse, autotune != 0));
                }
                catch (...)
                {
                    delete b;
                    throw;
                }
                return reinterpret_cast<KFR_AMD_BATCH*>(b);
            },
            (KFR_AMD_BATCH*)nullptr);
    }
    size_t kfr_amd_batch_total_size(KFR_AMD_BATCH* batch)
    {
        return try_fn(
            [&]
            {
                check_plan(batch);
                return reinterpret_cast<batch_obj*>(batch)->total();
            },
            (size_t)0);
    }
    int kfr_amd_batch_execute(KFR_AMD_BATCH* batch, void* out, const void* in, void* stream)
    {
        return try_fn(
            [&]
            {
                check_plan(batch);
                auto* b = reinterpret_cast<batch_obj*>(batch);
                std::lock_guard<std::mutex> lock(b->mu);
                if (is_device_ptr(in) && is_device_ptr(out))
                {
                    b->run(out, in, as_stream(stream));
                    return 0;
                }
                // host data: staged through one device buffer, synchronously
                const size_t bytes = b->total() * b->elem();
                hipStream_t s      = library_stream();
                b->buf.ensure(bytes);
                hip_check(hipMemcpyAsync(b->buf.ptr, in, bytes, hipMemcpyDefault, s), "batch input copy");
                b->run(b->buf.ptr, b->buf.ptr, s);
                hip_check(hipMemcpyAsync(out, b->buf.ptr, bytes, hipMemcpyDefault, s), "batch output copy");
                hip_check(hipStreamSynchronize(s), "batch execute");
                return 0;
            },
            -1);
    }
    int kfr_amd_batch_describe(KFR_AMD_BATCH* batch, char* buf, size_t buflen)
    {
        return try_fn(
            [&]
            {
                check_plan(batch);
                auto* b = reinterpret_cast<batch_obj*>(batch);
                return copy_text(b->f ? b->f->describe() : b->d->describe(), buf, buflen);
            },
            -1);
    }
    void kfr_amd_batch_delete(KFR_AMD_BATCH* batch)
    {
        try_void([&] { delete reinterpret_cast<batch_obj*>(batch); });
    }

    KFR_AMD_DFT_HALF* kfr_amd_dft_half_create(int storage, size_t size)
    {
        return try_fn(
            [&]
            {
                if (storage != 1 && storage != 2)
                    throw std::invalid_argument("storage must be 1 (bf16) or 2 (f16)");
                if (size < 2 || (size & (size - 1)) || !pow2_fft<float>::supports(size))
                    throw std::invalid_argument("half-precision transforms take a power-of-two size from 2 to 2^28");
                auto* h = new half_obj();
                try
                {
                    h->n = size;
                    h->f.reset(new pow2_fft<float>(size, library_stream(), storage));
                    hip_check(hipStreamSynchronize(library_stream()), "plan initialisation");
                }
                catch (...)
                {
                    delete h;
                    throw;
                }
                return reinterpret_cast<KFR_AMD_DFT_HALF*>(h);
            },
            (KFR_AMD_DFT_HALF*)nullptr);
    }
    int kfr_amd_dft_half_execute_batch(KFR_AMD_DFT_HALF* plan, void* out, const void* in, size_t batch, int inverse,
                                       void* stream)
    {
        return try_fn(
            [&]
            {
                check_plan(plan);
                auto* h = reinterpret_cast<half_obj*>(plan);
                std::lock_guard<std::mutex> lock(h->mu);
                if (is_device_ptr(in) && is_device_ptr(out))
                {
                    h->run(inverse != 0, in, out, (long long)batch, as_stream(stream));
                    return 0;
                }
                const size_t bytes = batch * h->n * 4;
                hipStream_t s      = library_stream();
                h->buf.ensure(bytes);
                hip_check(hipMemcpyAsync(h->buf.ptr, in, bytes, hipMemcpyDefault, s), "input copy");
                h->run(inverse != 0, h->buf.ptr, h->buf.ptr, (long long)batch, s);
                hip_check(hipMemcpyAsync(out, h->buf.ptr, bytes, hipMemcpyDefault, s), "output copy");
                hip_check(hipStreamSynchronize(s), "execute");
                return 0;
            },
            -1);
    }
    int kfr_amd_dft_half_describe(KFR_AMD_DFT_HALF* plan, char* buf, size_t buflen)
    {
        return try_fn(
            [&]
            {
                check_plan(plan);
                return copy_text("gpu(gfx950): " + reinterpret_cast<half_obj*>(plan)->f->describe(), buf, buflen);
            },
            -1);
    }
    void kfr_amd_dft_half_delete(KFR_AMD_DFT_HALF* plan)
    {
        try_void([&] { delete reinterpret_cast<half_obj*>(plan); });
    }

    void kfr_amd_inject_alloc_failure(int n) { set_alloc_failure_countdown(n); }

    int kfr_amd_timing_begin(void)
    {
        return try_fn(
            []
            {
                timing::state& st = timing::get();
                std::lock_guard<std::mutex> g(st.m);
                for (auto& r : st.recs)
                {
                    st.pool.push_back(r.start);
                    st.pool.push_back(r.stop);
                }
                st.recs.clear();
                timing_results().clear();
                st.on = true;
                return 0;
            },
            -1);
    }

    long long kfr_amd_timing_end(void)
    {
        return try_fn(
            []
            {
                timing::state& st = timing::get();
                st.on = false;
                std::lock_guard<std::mutex> g(st.m);
                auto& out = timing_results();
                out.clear();
                std::string err;
                for (auto& r : st.recs)
                {
                    float ms      = 0.f;
                    hipError_t e  = hipEventSynchronize(r.stop);
                    if (e == hipSuccess)
                        e = hipEventElapsedTime(&ms, r.start, r.stop);
                    if (e != hipSuccess && err.empty())
                        err = hipGetErrorString(e);
                    out.push_back(timing_result{ r.name, r.bytes, (double)ms });
                    st.pool.push_back(r.start);
                    st.pool.push_back(r.stop);
                }
                st.recs.clear();
                if (!err.empty())
                    throw std::runtime_error("kernel timing: " + err);
                return (long long)out.size();
            },
            -1ll);
    }

    int kfr_amd_timing_get(long long index, char* name, size_t namelen, long long* bytes, double* ms)
    {
        return try_fn(
            [&]
            {
                const auto& out = timing_results();
                if (index < 0 || index >= (long long)out.size())
                    throw std::out_of_range("timing record index out of range");
                const timing_result& r = out[(size_t)index];
                if (name && namelen)
                {
                    const size_t n = std::min(namelen - 1, r.name.size());
                    std::memcpy(name, r.name.data(), n);
                    name[n] = 0;
                }
                if (bytes)
                    *bytes = r.bytes;
                if (ms)
                    *ms = r.ms;
                return 0;
            },
            -1);
    }

    // ---- plan cache
    KFR_DFT_PLAN_F32* kfr_amd_cache_dft_plan_f32(size_t size)
    {
        return try_fn(
            [&]
            {
                cache_key k;
                k.dtype = 0;
                k.n     = size;
                return reinterpret_cast<KFR_DFT_PLAN_F32*>(plan_cache::get().acquire<dft_obj<float>>(
                    k, [&] { return new dft_obj<float>(std::vector<size_t>{ size }); }));
            },
            (KFR_DFT_PLAN_F32*)nullptr);
    }
    KFR_DFT_PLAN_F64* kfr_amd_cache_dft_plan_f64(size_t size)
    {
        return try_fn(
            [&]
            {
                cache_key k;
                k.dtype = 1;
                k.n     = size;
                return reinterpret_cast<KFR_DFT_PLAN_F64*>(plan_cache::get().acquire<dft_obj<double>>(
                    k, [&] { return new dft_obj<double>(std::vector<size_t>{ size }); }));
            },
            (KFR_DFT_PLAN_F64*)nullptr);
    }
    KFR_DFT_REAL_PLAN_F32* kfr_amd_cache_dft_real_plan_f32(size_t size, KFR_DFT_PACK_FORMAT fmt)
    {
        return try_fn(
            [&]
            {
                cache_key k;
                k.dtype = 0;
                k.kind  = 1;
                k.n     = size;
                k.fmt   = (int)fmt;
                return reinterpret_cast<KFR_DFT_REAL_PLAN_F32*>(plan_cache::get().acquire<real_obj<float>>(
                    k, [&] { return new real_obj<float>(size, (int)fmt); }));
            },
            (KFR_DFT_REAL_PLAN_F32*)nullptr);
    }
    KFR_DFT_REAL_PLAN_F64* kfr_amd_cache_dft_real_plan_f64(size_t size, KFR_DFT_PACK_FORMAT fmt)
    {
        return try_fn(
            [&]
            {
                cache_key k;
                k.dtype = 1;
                k.kind  = 1;
                k.n     = size;
                k.fmt   = (int)fmt;
                return reinterpret_cast<KFR_DFT_REAL_PLAN_F64*>(plan_cache::get().acquire<real_obj<double>>(
                    k, [&] { return new real_obj<double>(size, (int)fmt); }));
            },
            (KFR_DFT_REAL_PLAN_F64*)nullptr);
    }
    void kfr_amd_cache_clear(void)
    {
        try_void([] { plan_cache::get().clear(); });
    }
    size_t kfr_amd_cache_size(void)
    {
        return try_fn([] { return plan_cache::get().size(); }, (size_t)0);
    }

    // ---- complex streaming filters
    KFR_FILTER_C32* kfr_amd_filter_create_convolution_plan_c32(const kfr_c32* taps, size_t size, size_t block_size)
    {
        return try_fn(
            [&] { return reinterpret_cast<KFR_FILTER_C32*>(new cfilter_obj<float>(taps, size, block_size ? block_size : 1024)); },
            (KFR_FILTER_C32*)nullptr);
    }
    KFR_FILTER_C64* kfr_amd_filter_create_convolution_plan_c64(const kfr_c64* taps, size_t size, size_t block_size)
    {
        return try_fn(
            [&]
            { return reinterpret_cast<KFR_FILTER_C64*>(new cfilter_obj<double>(taps, size, block_size ? block_size : 1024)); },
            (KFR_FILTER_C64*)nullptr);
    }
    void kfr_amd_filter_process_c32(KFR_FILTER_C32* plan, kfr_c32* output, const kfr_c32* input, size_t size)
    {
        try_void([&] { cfilter_process<float>(plan, output, input, size); });
    }
    void kfr_amd_filter_process_c64(KFR_FILTER_C64* plan, kfr_c64* output, const kfr_c64* input, size_t size)
    {
        try_void([&] { cfilter_process<double>(plan, output, input, size); });
    }
    void kfr_amd_filter_reset_c32(KFR_FILTER_C32* plan)
    {
        try_void(
            [&]
            {
                check_plan(plan);
                reinterpret_cast<cfilter_obj<float>*>(plan)->g.reset();
            });
    }
    void kfr_amd_filter_reset_c64(KFR_FILTER_C64* plan)
    {
        try_void(
            [&]
            {
                check_plan(plan);
                reinterpret_cast<cfilter_obj<double>*>(plan)->g.reset();
            });
    }
    void kfr_amd_filter_delete_plan_c32(KFR_FILTER_C32* plan)
    {
        try_void([&] { delete reinterpret_cast<cfilter_obj<float>*>(plan); });
    }
    void kfr_amd_filter_delete_plan_c64(KFR_FILTER_C64* plan)
    {
        try_void([&] { delete reinterpret_cast<cfilter_obj<double>*>(plan); });
    }
    void kfr_amd_filter_set_data_f32(KFR_FILTER_F32* plan, const kfr_f32* taps, size_t size)
    {
        try_void(
            [&]
            {
                check_plan(plan);
                auto* f = reinterpret_cast<filter_obj<float>*>(plan);
                std::lock_guard<std::mutex> lock(f->mu);
                f->g.set_data(taps, size);
            });
    }
    void kfr_amd_filter_set_data_f64(KFR_FILTER_F64* plan, const kfr_f64* taps, size_t size)
    {
        try_void(
            [&]
            {
                check_plan(plan);
                auto* f = reinterpret_cast<filter_obj<double>*>(plan);
                std::lock_guard<std::mutex> lock(f->mu);
                f->g.set_data(taps, size);
            });
    }
    void kfr_amd_filter_set_data_c32(KFR_FILTER_C32* plan, const kfr_c32* taps, size_t size)
    {
        try_void(
            [&]
            {
                check_plan(plan);
                auto* f = reinterpret_cast<cfilter_obj<float>*>(plan);
                std::lock_guard<std::mutex> lock(f->mu);
                f->g.set_data(taps, size);
            });
    }
    void kfr_amd_filter_set_data_c64(KFR_FILTER_C64* plan, const kfr_c64* taps, size_t size)
    {
        try_void(
            [&]
            {
                check_plan(plan);
                auto* f = reinterpret_cast<cfilter_obj<double>*>(plan);
                std::lock_guard<std::mutex> lock(f->mu);
                f->g.set_data(taps, size);
            });
    }

    // ---- progressive execution
    KFR_AMD_PROGRESSIVE* kfr_amd_dft_progressive_start_f32(KFR_DFT_PLAN_F32* plan, int progressive_optimized,
                                                           int inverse, kfr_c32* out, const kfr_c32* in)
    {
        return try_fn(
            [&]
            {
                return reinterpret_cast<KFR_AMD_PROGRESSIVE*>(static_cast<progressive_obj*>(
                    new dft_progressive<float>(dft_cast<float>(plan), progressive_optimized != 0, inverse != 0, out, in)));
            },
            (KFR_AMD_PROGRESSIVE*)nullptr);
    }
    KFR_AMD_PROGRESSIVE* kfr_amd_dft_progressive_start_f64(KFR_DFT_PLAN_F64* plan, int progressive_optimized,
                                                           int inverse, kfr_c64* out, const kfr_c64* in)
    {
        return try_fn(
            [&]
            {
                return reinterpret_cast<KFR_AMD_PROGRESSIVE*>(static_cast<progressive_obj*>(new dft_progressive<double>(
                    dft_cast<double>(plan), progressive_optimized != 0, inverse != 0, out, in)));
            },
            (KFR_AMD_PROGRESSIVE*)nullptr);
    }
    KFR_AMD_PROGRESSIVE* kfr_amd_dft_real_progressive_start_f32(KFR_DFT_REAL_PLAN_F32* plan, int inverse, void* out,
                                                                const void* in)
    {
        return try_fn(
            [&]
            {
                return reinterpret_cast<KFR_AMD_PROGRESSIVE*>(static_cast<progressive_obj*>(
                    new real_progressive<float>(real_cast<float>(plan), inverse != 0, out, in)));
            },
            (KFR_AMD_PROGRESSIVE*)nullptr);
    }
    KFR_AMD_PROGRESSIVE* kfr_amd_dft_real_progressive_start_f64(KFR_DFT_REAL_PLAN_F64* plan, int inverse, void* out,
                                                                const void* in)
    {
        return try_fn(
            [&]
            {
                return reinterpret_cast<KFR_AMD_PROGRESSIVE*>(static_cast<progressive_obj*>(
                    new real_progressive<double>(real_cast<double>(plan), inverse != 0, out, in)));
            },
            (KFR_AMD_PROGRESSIVE*)nullptr);
    }
    size_t kfr_amd_progressive_total_steps(KFR_AMD_PROGRESSIVE* prog)
    {
        return try_fn(
            [&]
            {
                check_plan(prog);
                return reinterpret_cast<progressive_obj*>(prog)->total;
            },
            (size_t)0);
    }
    int kfr_amd_progressive_step(KFR_AMD_PROGRESSIVE* prog)
    {
        return try_fn(
            [&]
            {
                check_plan(prog);
                auto* p = reinterpret_cast<progressive_obj*>(prog);
                if (p->step >= p->total)
                    throw std::logic_error("progressive execution is already complete");
                p->run_step(p->step);
                return ++p->step < p->total ? 1 : 0;
            },
            -1);
    }
    void kfr_amd_progressive_delete(KFR_AMD_PROGRESSIVE* prog)
    {
        try_void([&] { delete reinterpret_cast<progressive_obj*>(prog); });
    }
    size_t kfr_amd_dft_progressive_steps(int dtype, size_t size, int progressive_optimized)
    {
        return try_fn([&] { return describe_dft(to_dtype(dtype), size, progressive_optimized != 0).order[0].size(); },
                      (size_t)0);
    }
    size_t kfr_amd_dft_real_progressive_steps(int dtype, size_t size, KFR_DFT_PACK_FORMAT fmt)
    {
        return try_fn([&] { return describe_dft_real(to_dtype(dtype), size, (int)fmt).order[0].size(); }, (size_t)0);
    }

    int kfr_amd_synchronize(void)
    {
        return try_fn(
            []
            {
                hip_check(hipStreamSynchronize(library_stream()), "synchronize");
                return 0;
            },
            -1);
    }

} // extern "C"
