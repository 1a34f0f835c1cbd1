// Test infrastructure only: a C-ABI shim over the reference KFR DFT engine, compiled from the
// reference sources by oracle/Makefile into oracle/_ref/. It is loaded by tests/ (as the checker)
// and by bench.py's cpu_baseline leg (as the timed CPU baseline). Nothing in the product links it.
//
// Every entry point calls KFR's public C++ API:
//   dft_plan<T>                 include/kfr/dft/fft.hpp:183-456
//   dft_plan_real<T>            include/kfr/dft/fft.hpp:459-566
//   dct_plan<T>                 include/kfr/dft/fft.hpp:927-978
//   dft_plan_md / _md_real      include/kfr/dft/fft.hpp:569-924
//   convolve/correlate          include/kfr/dft/convolution.hpp:53-76
//   convolve_filter<T>          include/kfr/dft/convolution.hpp:99-148
// The plan description walks dft_plan::all_stages, the same fields dft_stage::dump() prints
// (include/kfr/dft/fft.hpp:77-81).
#include <kfr/dft.hpp>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include <unistd.h>

using namespace kfr;

#define KREF_API extern "C" __attribute__((visibility("default")))

namespace
{

template <typename T>
std::string describe_stages(const dft_plan<T>& plan)
{
    std::string s;
    char line[512];
    std::snprintf(line, sizeof(line), "size=%zu temp_size=%zu data_size=%zu stages=%zu,%zu arblen=%d\n",
                  plan.size, plan.temp_size, plan.data_size, plan.stages[0].size(), plan.stages[1].size(),
                  (int)plan.arblen);
    s += line;
    // the stage lines are the reference's own dump() output, captured from stdout
    {
        std::fflush(stdout);
        FILE* tf        = std::tmpfile();
        const int saved = dup(1);
        dup2(fileno(tf), 1);
        plan.dump();
        std::fflush(stdout);
        dup2(saved, 1);
        close(saved);
        std::rewind(tf);
        char buf[4096];
        size_t got;
        while ((got = std::fread(buf, 1, sizeof(buf), tf)) > 0)
            s.append(buf, got);
        std::fclose(tf);
    }
    // stage order of the forward and inverse lists, as indexes into all_stages
    for (int dir = 0; dir < 2; ++dir)
    {
        s += dir == 0 ? "fwd:" : "inv:";
        for (auto* p : plan.stages[dir])
        {
            for (size_t i = 0; i < plan.all_stages.size(); ++i)
                if (plan.all_stages[i].get() == p)
                {
                    std::snprintf(line, sizeof(line), " %zu", i);
                    s += line;
                }
        }
        s += "\n";
    }
    return s;
}

int copy_out(const std::string& s, char* buf, size_t buflen)
{
    if (buf && buflen > 0)
    {
        size_t n = std::min(buflen - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return (int)s.size();
}

template <typename F>
void parallel_batches(size_t batch, int threads, F&& fn)
{
    if (threads <= 1 || batch <= 1)
    {
        fn(size_t(0), batch, 0);
        return;
    }
    threads = (int)std::min<size_t>(threads, batch);
    std::vector<std::thread> pool;
    size_t per = (batch + threads - 1) / threads;
    for (int t = 0; t < threads; ++t)
    {
        size_t b0 = t * per, b1 = std::min(batch, b0 + per);
        if (b0 >= b1)
            break;
        pool.emplace_back([&, b0, b1, t] { fn(b0, b1, t); });
    }
    for (auto& th : pool)
        th.join();
}

template <typename T>
void run_dft(size_t n, size_t batch, void* out, const void* in, int inverse, int threads)
{
    dft_plan<T> plan(n);
    auto* o       = static_cast<complex<T>*>(out);
    const auto* i = static_cast<const complex<T>*>(in);
    parallel_batches(batch, threads,
                     [&](size_t b0, size_t b1, int)
                     {
                         std::vector<u8> temp(plan.temp_size + 64);
                         for (size_t b = b0; b < b1; ++b)
                             plan.execute(o + b * n, i + b * n, temp.data(), inverse != 0);
                     });
}

template <typename T>
void run_rdft(size_t n, int fmt, size_t batch, void* out, const void* in, int inverse, int threads)
{
    dft_plan_real<T> plan(n, fmt == 0 ? dft_pack_format::Perm : dft_pack_format::CCs);
    const size_t cs = plan.complex_size();
    parallel_batches(batch, threads,
                     [&](size_t b0, size_t b1, int)
                     {
                         std::vector<u8> temp(plan.temp_size + 64);
                         for (size_t b = b0; b < b1; ++b)
                         {
                             if (!inverse)
                                 plan.execute(static_cast<complex<T>*>(out) + b * cs,
                                              static_cast<const T*>(in) + b * n, temp.data(), cdirect_t{});
                             else
                                 plan.execute(static_cast<T*>(out) + b * n,
                                              static_cast<const complex<T>*>(in) + b * cs, temp.data(),
                                              cinvert_t{});
                         }
                     });
}

template <typename T>
void run_dct(size_t n, size_t batch, void* out, const void* in, int inverse, int threads)
{
    dct_plan<T> plan(n);
    parallel_batches(batch, threads,
                     [&](size_t b0, size_t b1, int)
                     {
                         std::vector<u8> temp(plan.temp_size + 64);
                         for (size_t b = b0; b < b1; ++b)
                             plan.execute(static_cast<T*>(out) + b * n, static_cast<const T*>(in) + b * n,
                                          temp.data(), inverse != 0);
                     });
}

template <typename T>
int find_reorder(const dft_plan<T>& plan)
{
    for (size_t i = 0; i < plan.all_stages.size(); ++i)
    {
        const char* nm = plan.all_stages[i]->name;
        if (nm && (std::strncmp(nm, "fft_reorder_stage_impl", 22) == 0 ||
                   std::strncmp(nm, "dft_reorder_stage_impl", 22) == 0))
            return (int)i;
    }
    return -1;
}

template <typename T>
int run_permutation(size_t n, int64_t* out_idx)
{
    dft_plan<T> plan(n);
    int si = find_reorder(plan);
    if (si < 0)
        return 0;
    std::vector<complex<T>> a(n), b(n);
    for (size_t i = 0; i < n; ++i)
        a[i] = complex<T>(T(i), T(0));
    std::vector<u8> temp(plan.temp_size + 64);
    auto* st = plan.all_stages[si].get();
    // fft_reorder_stage_impl permutes `out` in place and ignores `in` (src/dft/fft-impl.hpp:997-1000);
    // dft_reorder_stage_impl is out of place (src/dft/dft-impl.hpp:428-494).
    if (std::strncmp(st->name, "fft_reorder", 11) == 0)
    {
        b = a;
        st->execute(false, b.data(), b.data(), temp.data());
    }
    else
        st->execute(false, b.data(), a.data(), temp.data());
    for (size_t i = 0; i < n; ++i)
        out_idx[i] = (int64_t)b[i].real();
    return 1;
}

struct filter_box
{
    int dtype;
    std::unique_ptr<convolve_filter<float>> f32;
    std::unique_ptr<convolve_filter<double>> f64;
    std::unique_ptr<convolve_filter<complex<float>>> c32;
    std::unique_ptr<convolve_filter<complex<double>>> c64;
};

} // namespace

KREF_API const char* kref_version() { return library_version_dft(); }

// kind: 0 = dft_plan, 1 = dft_plan_real (arg = pack format 0 Perm / 1 CCs), 2 = dct_plan,
//       3 = dft_plan with progressive_optimized = true
// dtype: 0 = float, 1 = double
KREF_API int kref_describe(int kind, int dtype, size_t n, int arg, char* buf, size_t buflen)
{
    try
    {
        std::string s;
        auto go = [&](auto tag)
        {
            using T = decltype(tag);
            if (kind == 0)
                s = describe_stages(dft_plan<T>(n));
            else if (kind == 1)
                s = describe_stages<T>(dft_plan_real<T>(n, arg == 0 ? dft_pack_format::Perm
                                                                    : dft_pack_format::CCs));
            else if (kind == 2)
                s = describe_stages<T>(dct_plan<T>(n));
            else
                s = describe_stages(dft_plan<T>(n, dft_order::normal, true));
        };
        if (dtype == 0)
            go(float{});
        else
            go(double{});
        return copy_out(s, buf, buflen);
    }
    catch (const std::exception& e)
    {
        return copy_out(std::string("error: ") + e.what(), buf, buflen), -1;
    }
}

KREF_API int kref_permutation(int dtype, size_t n, int64_t* out_idx)
{
    return dtype == 0 ? run_permutation<float>(n, out_idx) : run_permutation<double>(n, out_idx);
}

KREF_API int kref_dft(int dtype, size_t n, size_t batch, void* out, const void* in, int inverse, int threads)
{
    try
    {
        if (dtype == 0)
            run_dft<float>(n, batch, out, in, inverse, threads);
        else
            run_dft<double>(n, batch, out, in, inverse, threads);
        return 0;
    }
    catch (...)
    {
        return -1;
    }
}

// dft_plan_md / dft_plan_md_real (include/kfr/dft/fft.hpp:569-924) on one array; returns the plan's
// temp_size, or -1. The real inverse with real_out_is_enough needs `out` to hold real_out_size() values.
template <typename T>
long run_dft_md(size_t dims, const unsigned* shp, void* out, const void* in, int inverse)
{
    shape<dynamic_shape> sh(dims);
    for (size_t i = 0; i < dims; ++i)
        sh[i] = shp[i];
    dft_plan_md<T, dynamic_shape> plan(sh);
    std::vector<u8> temp(plan.temp_size + 64);
    plan.execute(static_cast<complex<T>*>(out), static_cast<const complex<T>*>(in), temp.data(), inverse != 0);
    return (long)plan.temp_size;
}
template <typename T>
long run_rdft_md(size_t dims, const unsigned* shp, int real_out_is_enough, void* out, const void* in, int inverse)
{
    shape<dynamic_shape> sh(dims);
    for (size_t i = 0; i < dims; ++i)
        sh[i] = shp[i];
    dft_plan_md_real<T, dynamic_shape> plan(sh, real_out_is_enough != 0);
    std::vector<u8> temp(plan.temp_size + 64);
    if (inverse)
        plan.execute(static_cast<T*>(out), static_cast<const complex<T>*>(in), temp.data(), cinvert_t{});
    else
        plan.execute(static_cast<complex<T>*>(out), static_cast<const T*>(in), temp.data(), cdirect_t{});
    return (long)plan.temp_size;
}

KREF_API long kref_dft_md(int dtype, size_t dims, const unsigned* shape, void* out, const void* in, int inverse)
{
    try
    {
        return dtype == 0 ? run_dft_md<float>(dims, shape, out, in, inverse)
                          : run_dft_md<double>(dims, shape, out, in, inverse);
    }
    catch (...)
    {
        return -1;
    }
}

KREF_API long kref_rdft_md(int dtype, size_t dims, const unsigned* shape, int real_out_is_enough, void* out,
                           const void* in, int inverse)
{
    try
    {
        return dtype == 0 ? run_rdft_md<float>(dims, shape, real_out_is_enough, out, in, inverse)
                          : run_rdft_md<double>(dims, shape, real_out_is_enough, out, in, inverse);
    }
    catch (...)
    {
        return -1;
    }
}

KREF_API int kref_rdft(int dtype, size_t n, int fmt, size_t batch, void* out, const void* in, int inverse,
                       int threads)
{
    try
    {
        if (dtype == 0)
            run_rdft<float>(n, fmt, batch, out, in, inverse, threads);
        else
            run_rdft<double>(n, fmt, batch, out, in, inverse, threads);
        return 0;
    }
    catch (...)
    {
        return -1;
    }
}

KREF_API int kref_dct(int dtype, size_t n, size_t batch, void* out, const void* in, int inverse, int threads)
{
    try
    {
        if (dtype == 0)
            run_dct<float>(n, batch, out, in, inverse, threads);
        else
            run_dct<double>(n, batch, out, in, inverse, threads);
        return 0;
    }
    catch (...)
    {
        return -1;
    }
}

// dtype: 0 f32, 1 f64, 2 c32, 3 c64. mode: 0 convolve, 1 correlate, 2 autocorrelate (b ignored).
// Returns the output length; writes it only when out != NULL.
KREF_API long kref_convolve(int dtype, int mode, const void* a, size_t na, const void* b, size_t nb, void* out)
{
    auto go = [&](auto tag) -> long
    {
        using T = decltype(tag);
        univector<T> x(static_cast<const T*>(a), static_cast<const T*>(a) + na);
        univector<T> y;
        if (mode != 2)
            y = univector<T>(static_cast<const T*>(b), static_cast<const T*>(b) + nb);
        univector<T> r = mode == 0 ? convolve(x, y) : mode == 1 ? correlate(x, y) : autocorrelate(x);
        if (out)
            std::copy(r.begin(), r.end(), static_cast<T*>(out));
        return (long)r.size();
    };
    switch (dtype)
    {
    case 0:
        return go(float{});
    case 1:
        return go(double{});
    case 2:
        return go(complex<float>{});
    default:
        return go(complex<double>{});
    }
}

KREF_API void* kref_filter_create(int dtype, const void* taps, size_t ntaps, size_t block)
{
    auto* box  = new filter_box();
    box->dtype = dtype;
    switch (dtype)
    {
    case 0:
        box->f32.reset(new convolve_filter<float>(
            univector_ref<const float>(static_cast<const float*>(taps), ntaps), block));
        break;
    case 1:
        box->f64.reset(new convolve_filter<double>(
            univector_ref<const double>(static_cast<const double*>(taps), ntaps), block));
        break;
    case 2:
        box->c32.reset(new convolve_filter<complex<float>>(
            univector_ref<const complex<float>>(static_cast<const complex<float>*>(taps), ntaps), block));
        break;
    default:
        box->c64.reset(new convolve_filter<complex<double>>(
            univector_ref<const complex<double>>(static_cast<const complex<double>*>(taps), ntaps), block));
    }
    return box;
}

KREF_API void kref_filter_apply(void* h, void* out, const void* in, size_t n)
{
    auto* box = static_cast<filter_box*>(h);
    switch (box->dtype)
    {
    case 0:
        box->f32->apply(static_cast<float*>(out), static_cast<const float*>(in), n);
        break;
    case 1:
        box->f64->apply(static_cast<double*>(out), static_cast<const double*>(in), n);
        break;
    case 2:
        box->c32->apply(static_cast<complex<float>*>(out), static_cast<const complex<float>*>(in), n);
        break;
    default:
        box->c64->apply(static_cast<complex<double>*>(out), static_cast<const complex<double>*>(in), n);
    }
}

KREF_API void kref_filter_reset(void* h)
{
    auto* box = static_cast<filter_box*>(h);
    if (box->f32)
        box->f32->reset();
    if (box->f64)
        box->f64->reset();
    if (box->c32)
        box->c32->reset();
    if (box->c64)
        box->c64->reset();
}

KREF_API void kref_filter_delete(void* h) { delete static_cast<filter_box*>(h); }

// CPU baseline timing: one shared plan, the batch split over `threads` threads, each with its own temp
// buffer (the protocol of tests/dft_test.cpp:56-81 with planning excluded). kind 0 = C2C, 1 = R2C (CCs),
// 2 = DCT-II. Runs the whole batch `reps` times and returns the median seconds per batch.
KREF_API double kref_time(int kind, int dtype, size_t n, size_t batch, int inverse, int threads, int reps,
                          void* out, const void* in)
{
    auto go = [&](auto tag) -> double
    {
        using T = decltype(tag);
        std::unique_ptr<dft_plan<T>> plan;
        std::unique_ptr<dft_plan_real<T>> rplan;
        std::unique_ptr<dct_plan<T>> cplan;
        size_t temp_size;
        if (kind == 0)
        {
            plan.reset(new dft_plan<T>(n));
            temp_size = plan->temp_size;
        }
        else if (kind == 1)
        {
            rplan.reset(new dft_plan_real<T>(n, dft_pack_format::CCs));
            temp_size = rplan->temp_size;
        }
        else
        {
            cplan.reset(new dct_plan<T>(n));
            temp_size = cplan->temp_size;
        }
        const size_t cs = n / 2 + 1;
        std::vector<double> times;
        for (int r = 0; r < reps; ++r)
        {
            auto t0 = std::chrono::steady_clock::now();
            parallel_batches(batch, threads,
                             [&](size_t b0, size_t b1, int)
                             {
                                 std::vector<u8> temp(temp_size + 64);
                                 for (size_t b = b0; b < b1; ++b)
                                 {
                                     if (kind == 0)
                                         plan->execute(static_cast<complex<T>*>(out) + b * n,
                                                       static_cast<const complex<T>*>(in) + b * n,
                                                       temp.data(), inverse != 0);
                                     else if (kind == 1)
                                         rplan->execute(static_cast<complex<T>*>(out) + b * cs,
                                                        static_cast<const T*>(in) + b * n, temp.data(),
                                                        cdirect_t{});
                                     else
                                         cplan->execute(static_cast<T*>(out) + b * n,
                                                        static_cast<const T*>(in) + b * n, temp.data(),
                                                        inverse != 0);
                                 }
                             });
            auto t1 = std::chrono::steady_clock::now();
            times.push_back(std::chrono::duration<double>(t1 - t0).count());
        }
        std::sort(times.begin(), times.end());
        return times[times.size() / 2];
    };
    return dtype == 0 ? go(float{}) : go(double{});
}
