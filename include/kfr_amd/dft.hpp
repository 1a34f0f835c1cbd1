// The reference's C++ DFT class API (namespace kfr) over the gfx950 engine's C ABI: a program written
// against kfr::dft_plan<T>, dft_plan_real<T>, dct_plan<T>, dft_plan_md<T>, dft_plan_md_real<T>,
// convolve / correlate / autocorrelate, convolve_filter<T> and dft_cache compiles against this header
// and links libkfr_amd.so. Names, argument meaning, buffer sizes and error behaviour follow
//   include/kfr/dft/fft.hpp:110-978        (dft_order, dft_pack_format, the plan classes)
//   include/kfr/dft/convolution.hpp:53-148 (convolve, correlate, autocorrelate, convolve_filter)
//   include/kfr/dft/cache.hpp:44-170       (dft_cache, dft, idft, realdft, irealdft)
// Buffers are host or device pointers; every call is synchronous, as in the reference. Errors that the
// reference reports by throwing (its KFR_LOGIC_CHECK and allocation failures) throw kfr::logic_error
// with the library's message. univector<T> here is the reference's dynamic univector reduced to what
// the DFT API uses: a contiguous, 64-byte aligned std::vector (data(), size(), element access).
// Beyond the reference: execute_batch() on device buffers and HIP streams.
#pragma once

#include "../kfr_amd_capi.h"

#include <array>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <limits>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace kfr
{

using u8    = uint8_t;
using f32   = float;
using f64   = double;
using index_t = size_t;
template <typename T>
using complex = std::complex<T>;

template <bool B>
using cbool_t                    = std::integral_constant<bool, B>;
using ctrue_t                    = cbool_t<true>;
using cfalse_t                   = cbool_t<false>;
using cdirect_t                  = cfalse_t;
using cinvert_t                  = ctrue_t;
constexpr inline ctrue_t ctrue   = {};
constexpr inline cfalse_t cfalse = {};
template <typename T>
struct ctype_t
{
};
template <typename T>
constexpr inline ctype_t<T> ctype = {};

struct logic_error : std::logic_error
{
    using std::logic_error::logic_error;
};

// ---- univector (include/kfr/base/univector.hpp): the dynamic form and the non-owning reference
template <typename T>
struct aligned_allocator
{
    using value_type = T;
    aligned_allocator() = default;
    template <typename U>
    aligned_allocator(const aligned_allocator<U>&)
    {
    }
    T* allocate(size_t n)
    {
        void* p = kfr_allocate_aligned(n > 0 ? n * sizeof(T) : 1, 64);
        if (!p)
            throw std::bad_alloc();
        return static_cast<T*>(p);
    }
    void deallocate(T* p, size_t) { kfr_deallocate(p); }
    template <typename U>
    bool operator==(const aligned_allocator<U>&) const
    {
        return true;
    }
    template <typename U>
    bool operator!=(const aligned_allocator<U>&) const
    {
        return false;
    }
};

using univector_tag                        = size_t;
constexpr inline univector_tag tag_dynamic_vector = std::numeric_limits<size_t>::max();
constexpr inline univector_tag tag_array_ref      = std::numeric_limits<size_t>::max() - 1;

// fixed size: Tag is the element count
template <typename T, univector_tag Tag = tag_dynamic_vector>
struct univector : std::array<T, Tag>
{
    univector() : std::array<T, Tag>{} {}
    univector(std::initializer_list<T> il) : std::array<T, Tag>{}
    {
        size_t i = 0;
        for (const T& v : il)
            if (i < Tag)
                (*this)[i++] = v;
    }
};

template <typename T>
struct univector<T, tag_dynamic_vector> : std::vector<T, aligned_allocator<T>>
{
    using base = std::vector<T, aligned_allocator<T>>;
    using base::base;
    univector() = default;
    univector(std::initializer_list<T> il) : base(il) {}
    univector<T, tag_array_ref> slice(size_t start = 0, size_t size = std::numeric_limits<size_t>::max());
    univector<const T, tag_array_ref> slice(size_t start = 0, size_t size = std::numeric_limits<size_t>::max()) const;
};

template <typename T>
struct univector<T, tag_array_ref>
{
    T* ptr      = nullptr;
    size_t count = 0;
    univector() = default;
    univector(T* p, size_t n) : ptr(p), count(n) {}
    template <univector_tag Tag>
    univector(univector<std::remove_const_t<T>, Tag>& v) : ptr(v.data()), count(v.size())
    {
    }
    template <univector_tag Tag>
    univector(const univector<std::remove_const_t<T>, Tag>& v) : ptr(v.data()), count(v.size())
    {
    }
    T* data() const { return ptr; }
    size_t size() const { return count; }
    T& operator[](size_t i) const { return ptr[i]; }
    T* begin() const { return ptr; }
    T* end() const { return ptr + count; }
    univector<T, tag_array_ref> slice(size_t start = 0, size_t size = std::numeric_limits<size_t>::max()) const
    {
        start = std::min(start, count);
        return { ptr + start, std::min(size, count - start) };
    }
};

template <typename T>
using univector_ref = univector<T, tag_array_ref>;

template <typename T>
univector<T, tag_array_ref> univector<T, tag_dynamic_vector>::slice(size_t start, size_t size)
{
    start = std::min(start, this->size());
    return { this->data() + start, std::min(size, this->size() - start) };
}
template <typename T>
univector<const T, tag_array_ref> univector<T, tag_dynamic_vector>::slice(size_t start, size_t size) const
{
    start = std::min(start, this->size());
    return { this->data() + start, std::min(size, this->size() - start) };
}

template <typename T>
univector_ref<T> make_univector(T* data, size_t size)
{
    return { data, size };
}

// ---- enums (include/kfr/dft/fft.hpp:110-140)
enum class dft_type
{
    both,
    direct,
    inverse
};
enum class dft_order
{
    normal,
    internal,
};
enum class dft_pack_format
{
    Perm,
    CCs,
};

namespace amd
{
// The library reports failures through kfr_last_error(); the class API throws them.
inline void check(bool ok = true)
{
    const char* e = kfr_last_error();
    if (e && e[0])
        throw logic_error(e);
    if (!ok)
        throw logic_error("kfr_amd call failed");
}

// the C ABI of one element type
template <typename T>
struct capi;
template <>
struct capi<float>
{
    using plan = KFR_DFT_PLAN_F32;
    using real_plan = KFR_DFT_REAL_PLAN_F32;
    using dct_plan  = KFR_DCT_PLAN_F32;
    using filter    = KFR_FILTER_F32;
    using cfilter   = KFR_FILTER_C32;
    using cx        = kfr_c32;
    static constexpr int dtype = 0;
    static plan* create(size_t n) { return kfr_dft_create_plan_f32(n); }
    static plan* create_md(size_t dims, const unsigned* shape) { return kfr_dft_create_md_plan_f32(dims, shape); }
    static plan* cached(size_t n) { return kfr_amd_cache_dft_plan_f32(n); }
    static void del(plan* p) { kfr_dft_delete_plan_f32(p); }
    static void dump(plan* p) { kfr_dft_dump_f32(p); }
    static size_t size(plan* p) { return kfr_dft_get_size_f32(p); }
    static size_t temp(plan* p) { return kfr_dft_get_temp_size_f32(p); }
    static void exec(plan* p, void* o, const void* i, u8* t, bool inv)
    {
        if (inv)
            kfr_dft_execute_inverse_f32(p, (cx*)o, (const cx*)i, t);
        else
            kfr_dft_execute_f32(p, (cx*)o, (const cx*)i, t);
    }
    static int exec_batch(plan* p, void* o, const void* i, size_t b, bool inv, void* s)
    {
        return kfr_amd_dft_execute_batch_f32(p, (cx*)o, (const cx*)i, b, inv, s);
    }
    static KFR_AMD_PROGRESSIVE* prog(plan* p, bool opt, bool inv, void* o, const void* i)
    {
        return kfr_amd_dft_progressive_start_f32(p, opt, inv, (cx*)o, (const cx*)i);
    }
    static real_plan* rcreate(size_t n, int fmt) { return kfr_dft_real_create_plan_f32(n, (KFR_DFT_PACK_FORMAT)fmt); }
    static real_plan* rcreate_md(size_t dims, const unsigned* shape, bool roe)
    {
        return kfr_dft_real_create_md_plan_f32(dims, shape, roe);
    }
    static real_plan* rcached(size_t n, int fmt) { return kfr_amd_cache_dft_real_plan_f32(n, (KFR_DFT_PACK_FORMAT)fmt); }
    static void rdel(real_plan* p) { kfr_dft_real_delete_plan_f32(p); }
    static void rdump(real_plan* p) { kfr_dft_real_dump_f32(p); }
    static size_t rtemp(real_plan* p) { return kfr_dft_real_get_temp_size_f32(p); }
    static void rfwd(real_plan* p, void* o, const void* i, u8* t) { kfr_dft_real_execute_f32(p, (cx*)o, (const float*)i, t); }
    static void rinv(real_plan* p, void* o, const void* i, u8* t)
    {
        kfr_dft_real_execute_inverse_f32(p, (float*)o, (const cx*)i, t);
    }
    static KFR_AMD_PROGRESSIVE* rprog(real_plan* p, bool inv, void* o, const void* i)
    {
        return kfr_amd_dft_real_progressive_start_f32(p, inv, o, i);
    }
    static dct_plan* dcreate(size_t n) { return kfr_dct_create_plan_f32(n); }
    static void ddel(dct_plan* p) { kfr_dct_delete_plan_f32(p); }
    static void ddump(dct_plan* p) { kfr_dct_dump_f32(p); }
    static size_t dtemp(dct_plan* p) { return kfr_dct_get_temp_size_f32(p); }
    static void dexec(dct_plan* p, float* o, const float* i, u8* t, bool inv)
    {
        if (inv)
            kfr_dct_execute_inverse_f32(p, o, i, t);
        else
            kfr_dct_execute_f32(p, o, i, t);
    }
    static filter* fcreate(const float* h, size_t n, size_t b) { return kfr_filter_create_convolution_plan_f32(h, n, b); }
    static void fproc(filter* f, float* o, const float* i, size_t n) { kfr_filter_process_f32(f, o, i, n); }
    static void freset(filter* f) { kfr_filter_reset_f32(f); }
    static void fset(filter* f, const float* h, size_t n) { kfr_amd_filter_set_data_f32(f, h, n); }
    static void fdel(filter* f) { kfr_filter_delete_plan_f32(f); }
    static cfilter* cfcreate(const void* h, size_t n, size_t b)
    {
        return kfr_amd_filter_create_convolution_plan_c32((const cx*)h, n, b);
    }
    static void cfproc(cfilter* f, void* o, const void* i, size_t n) { kfr_amd_filter_process_c32(f, (cx*)o, (const cx*)i, n); }
    static void cfreset(cfilter* f) { kfr_amd_filter_reset_c32(f); }
    static void cfset(cfilter* f, const void* h, size_t n) { kfr_amd_filter_set_data_c32(f, (const cx*)h, n); }
    static void cfdel(cfilter* f) { kfr_amd_filter_delete_plan_c32(f); }
};
template <>
struct capi<double>
{
    using plan = KFR_DFT_PLAN_F64;
    using real_plan = KFR_DFT_REAL_PLAN_F64;
    using dct_plan  = KFR_DCT_PLAN_F64;
    using filter    = KFR_FILTER_F64;
    using cfilter   = KFR_FILTER_C64;
    using cx        = kfr_c64;
    static constexpr int dtype = 1;
    static plan* create(size_t n) { return kfr_dft_create_plan_f64(n); }
    static plan* create_md(size_t dims, const unsigned* shape) { return kfr_dft_create_md_plan_f64(dims, shape); }
    static plan* cached(size_t n) { return kfr_amd_cache_dft_plan_f64(n); }
    static void del(plan* p) { kfr_dft_delete_plan_f64(p); }
    static void dump(plan* p) { kfr_dft_dump_f64(p); }
    static size_t size(plan* p) { return kfr_dft_get_size_f64(p); }
    static size_t temp(plan* p) { return kfr_dft_get_temp_size_f64(p); }
    static void exec(plan* p, void* o, const void* i, u8* t, bool inv)
    {
        if (inv)
            kfr_dft_execute_inverse_f64(p, (cx*)o, (const cx*)i, t);
        else
            kfr_dft_execute_f64(p, (cx*)o, (const cx*)i, t);
    }
    static int exec_batch(plan* p, void* o, const void* i, size_t b, bool inv, void* s)
    {
        return kfr_amd_dft_execute_batch_f64(p, (cx*)o, (const cx*)i, b, inv, s);
    }
    static KFR_AMD_PROGRESSIVE* prog(plan* p, bool opt, bool inv, void* o, const void* i)
    {
        return kfr_amd_dft_progressive_start_f64(p, opt, inv, (cx*)o, (const cx*)i);
    }
    static real_plan* rcreate(size_t n, int fmt) { return kfr_dft_real_create_plan_f64(n, (KFR_DFT_PACK_FORMAT)fmt); }
    static real_plan* rcreate_md(size_t dims, const unsigned* shape, bool roe)
    {
        return kfr_dft_real_create_md_plan_f64(dims, shape, roe);
    }
    static real_plan* rcached(size_t n, int fmt) { return kfr_amd_cache_dft_real_plan_f64(n, (KFR_DFT_PACK_FORMAT)fmt); }
    static void rdel(real_plan* p) { kfr_dft_real_delete_plan_f64(p); }
    static void rdump(real_plan* p) { kfr_dft_real_dump_f64(p); }
    static size_t rtemp(real_plan* p) { return kfr_dft_real_get_temp_size_f64(p); }
    static void rfwd(real_plan* p, void* o, const void* i, u8* t)
    {
        kfr_dft_real_execute_f64(p, (cx*)o, (const double*)i, t);
    }
    static void rinv(real_plan* p, void* o, const void* i, u8* t)
    {
        kfr_dft_real_execute_inverse_f64(p, (double*)o, (const cx*)i, t);
    }
    static KFR_AMD_PROGRESSIVE* rprog(real_plan* p, bool inv, void* o, const void* i)
    {
        return kfr_amd_dft_real_progressive_start_f64(p, inv, o, i);
    }
    static dct_plan* dcreate(size_t n) { return kfr_dct_create_plan_f64(n); }
    static void ddel(dct_plan* p) { kfr_dct_delete_plan_f64(p); }
    static void ddump(dct_plan* p) { kfr_dct_dump_f64(p); }
    static size_t dtemp(dct_plan* p) { return kfr_dct_get_temp_size_f64(p); }
    static void dexec(dct_plan* p, double* o, const double* i, u8* t, bool inv)
    {
        if (inv)
            kfr_dct_execute_inverse_f64(p, o, i, t);
        else
            kfr_dct_execute_f64(p, o, i, t);
    }
    static filter* fcreate(const double* h, size_t n, size_t b) { return kfr_filter_create_convolution_plan_f64(h, n, b); }
    static void fproc(filter* f, double* o, const double* i, size_t n) { kfr_filter_process_f64(f, o, i, n); }
    static void freset(filter* f) { kfr_filter_reset_f64(f); }
    static void fset(filter* f, const double* h, size_t n) { kfr_amd_filter_set_data_f64(f, h, n); }
    static void fdel(filter* f) { kfr_filter_delete_plan_f64(f); }
    static cfilter* cfcreate(const void* h, size_t n, size_t b)
    {
        return kfr_amd_filter_create_convolution_plan_c64((const cx*)h, n, b);
    }
    static void cfproc(cfilter* f, void* o, const void* i, size_t n) { kfr_amd_filter_process_c64(f, (cx*)o, (const cx*)i, n); }
    static void cfreset(cfilter* f) { kfr_amd_filter_reset_c64(f); }
    static void cfset(cfilter* f, const void* h, size_t n) { kfr_amd_filter_set_data_c64(f, (const cx*)h, n); }
    static void cfdel(cfilter* f) { kfr_amd_filter_delete_plan_c64(f); }
};

template <typename T>
constexpr void check_type()
{
    static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>, "T must be float or double");
}

// an owned library handle with its deleter
template <typename H, void (*Del)(H*)>
struct handle
{
    H* h = nullptr;
    handle() = default;
    explicit handle(H* p) : h(p) {}
    handle(const handle&) = delete;
    handle& operator=(const handle&) = delete;
    handle(handle&& o) noexcept : h(std::exchange(o.h, nullptr)) {}
    handle& operator=(handle&& o) noexcept
    {
        if (this != &o)
        {
            reset();
            h = std::exchange(o.h, nullptr);
        }
        return *this;
    }
    ~handle() { reset(); }
    void reset()
    {
        if (h)
            Del(h);
        h = nullptr;
    }
    H* get() const
    {
        if (!h)
            throw logic_error("the plan is not initialized");
        return h;
    }
};

// a progressive execution in flight (shared by copies of dft_plan<T>::progressive)
struct progressive_handle
{
    KFR_AMD_PROGRESSIVE* h = nullptr;
    explicit progressive_handle(KFR_AMD_PROGRESSIVE* p) : h(p) {}
    ~progressive_handle()
    {
        if (h)
            kfr_amd_progressive_delete(h);
    }
};
} // namespace amd

template <typename T>
struct dft_plan_real;
template <typename T>
struct dft_cache_impl_t;

/// Complex 1-D DFT (include/kfr/dft/fft.hpp:183-456)
template <typename T>
struct dft_plan
{
    using api = amd::capi<T>;
    /// The size of the DFT as passed to the constructor.
    size_t size = 0;
    /// The temporary (scratch) buffer size of the reference plan (accepted and unused by the GPU plan).
    size_t temp_size = 0;

    dft_plan() { amd::check_type<T>(); }
    dft_plan(const dft_plan&)            = delete;
    dft_plan& operator=(const dft_plan&) = delete;
    dft_plan(dft_plan&&)                 = default;
    dft_plan& operator=(dft_plan&&)      = default;

    explicit dft_plan(size_t size, dft_order order = dft_order::normal, bool progressive_optimized = false)
        : size(size), progressive_optimized(progressive_optimized), h(api::create(size))
    {
        (void)order;
        amd::check(h.h != nullptr);
        temp_size = api::temp(h.h);
    }

    bool is_initialized() const { return size != 0; }

    void dump() const { api::dump(h.get()); }

    void execute(complex<T>* out, const complex<T>* in, u8* temp, bool inverse = false) const
    {
        api::exec(h.get(), out, in, temp, inverse);
        amd::check();
    }
    template <bool inverse>
    void execute(complex<T>* out, const complex<T>* in, u8* temp, cbool_t<inverse>) const
    {
        execute(out, in, temp, inverse);
    }
    template <univector_tag Tag1, univector_tag Tag2, univector_tag Tag3>
    void execute(univector<complex<T>, Tag1>& out, const univector<complex<T>, Tag2>& in, univector<u8, Tag3>& temp,
                 bool inverse = false) const
    {
        execute(out.data(), in.data(), temp.data(), inverse);
    }
    template <bool inverse, univector_tag Tag1, univector_tag Tag2, univector_tag Tag3>
    void execute(univector<complex<T>, Tag1>& out, const univector<complex<T>, Tag2>& in, univector<u8, Tag3>& temp,
                 cbool_t<inverse>) const
    {
        execute(out.data(), in.data(), temp.data(), inverse);
    }
    template <univector_tag Tag1, univector_tag Tag2>
    void execute(univector<complex<T>, Tag1>& out, const univector<complex<T>, Tag2>& in, u8* temp,
                 bool inverse = false) const
    {
        execute(out.data(), in.data(), temp, inverse);
    }
    template <bool inverse, univector_tag Tag1, univector_tag Tag2>
    void execute(univector<complex<T>, Tag1>& out, const univector<complex<T>, Tag2>& in, u8* temp,
                 cbool_t<inverse>) const
    {
        execute(out.data(), in.data(), temp, inverse);
    }

    /// Extension: `batch` contiguous transforms of device buffers, enqueued on a HIP stream (NULL:
    /// the default stream) without synchronising.
    void execute_batch(complex<T>* out, const complex<T>* in, size_t batch, bool inverse = false,
                       void* stream = nullptr) const
    {
        amd::check(api::exec_batch(h.get(), out, in, batch, inverse, stream) == 0);
    }

    /// Progressive execution state (include/kfr/dft/fft.hpp:364-374).
    struct progressive
    {
        bool inverse = false;
        complex<T>* out = nullptr;
        const complex<T>* in = nullptr;
        u8* temp    = nullptr;
        size_t step = 0;
        std::shared_ptr<amd::progressive_handle> state;
    };

    size_t progressive_total_steps() const
    {
        const size_t n = kfr_amd_dft_progressive_steps(api::dtype, size, progressive_optimized);
        amd::check();
        return n;
    }

    progressive progressive_start(bool inverse, complex<T>* out, const complex<T>* in, u8* temp) const
    {
        if (!is_initialized())
            throw logic_error("dft_plan is not initialized");
        if (temp_size != 0 && temp == nullptr)
            throw logic_error("Temporary buffer must be provided for progressive execution");
        progressive r;
        r.inverse = inverse;
        r.out     = out;
        r.in      = in;
        r.temp    = temp;
        KFR_AMD_PROGRESSIVE* p = api::prog(h.get(), progressive_optimized, inverse, out, in);
        amd::check(p != nullptr);
        r.state = std::make_shared<amd::progressive_handle>(p);
        return r;
    }

    bool progressive_step(progressive& prog) const
    {
        if (!prog.state)
            throw logic_error("progressive execution was not started");
        const int r = kfr_amd_progressive_step(prog.state->h);
        amd::check(r >= 0);
        ++prog.step;
        return r == 1;
    }

    bool progressive_optimized = false;

    typename api::plan* native() const { return h.get(); }

protected:
    friend struct dft_cache_impl_t<T>;
    struct adopt
    {
    };
    dft_plan(adopt, typename api::plan* p) : h(p)
    {
        amd::check(p != nullptr);
        size      = api::size(p);
        temp_size = api::temp(p);
    }
    amd::handle<typename api::plan, &api::del> h;
};

/// Real-to-complex and complex-to-real 1-D DFT (include/kfr/dft/fft.hpp:459-566)
template <typename T>
struct dft_plan_real
{
    using api = amd::capi<T>;
    size_t size = 0;
    dft_pack_format fmt = dft_pack_format::CCs;
    size_t temp_size = 0;

    dft_plan_real() { amd::check_type<T>(); }
    dft_plan_real(const dft_plan_real&)            = delete;
    dft_plan_real& operator=(const dft_plan_real&) = delete;
    dft_plan_real(dft_plan_real&&)                 = default;
    dft_plan_real& operator=(dft_plan_real&&)      = default;

    explicit dft_plan_real(size_t size, dft_pack_format fmt = dft_pack_format::CCs, bool progressive_optimized = false)
        : size(size), fmt(fmt)
    {
        (void)progressive_optimized;
        if (size % 2 != 0)
            throw logic_error("dft_plan_real requires size to be even");
        h = decltype(h)(api::rcreate(size, (int)fmt));
        amd::check(h.h != nullptr);
        temp_size = api::rtemp(h.h);
    }

    bool is_initialized() const { return size != 0; }
    size_t complex_size() const { return complex_size_for(size, fmt); }
    constexpr static size_t complex_size_for(size_t size, dft_pack_format fmt)
    {
        return fmt == dft_pack_format::CCs ? size / 2 + 1 : size / 2;
    }
    void dump() const { api::rdump(h.get()); }

    void execute(complex<T>* out, const T* in, u8* temp, cdirect_t = {}) const
    {
        api::rfwd(h.get(), out, in, temp);
        amd::check();
    }
    void execute(T* out, const complex<T>* in, u8* temp, cinvert_t = {}) const
    {
        api::rinv(h.get(), out, in, temp);
        amd::check();
    }
    template <univector_tag Tag1, univector_tag Tag2, univector_tag Tag3>
    void execute(univector<complex<T>, Tag1>& out, const univector<T, Tag2>& in, univector<u8, Tag3>& temp,
                 cdirect_t = {}) const
    {
        execute(out.data(), in.data(), temp.data());
    }
    template <univector_tag Tag1, univector_tag Tag2, univector_tag Tag3>
    void execute(univector<T, Tag1>& out, const univector<complex<T>, Tag2>& in, univector<u8, Tag3>& temp,
                 cinvert_t = {}) const
    {
        execute(out.data(), in.data(), temp.data());
    }
    template <univector_tag Tag1, univector_tag Tag2>
    void execute(univector<complex<T>, Tag1>& out, const univector<T, Tag2>& in, u8* temp, cdirect_t = {}) const
    {
        execute(out.data(), in.data(), temp);
    }
    template <univector_tag Tag1, univector_tag Tag2>
    void execute(univector<T, Tag1>& out, const univector<complex<T>, Tag2>& in, u8* temp, cinvert_t = {}) const
    {
        execute(out.data(), in.data(), temp);
    }

    using progressive = typename dft_plan<T>::progressive;

    size_t progressive_total_steps() const
    {
        const size_t n = kfr_amd_dft_real_progressive_steps(api::dtype, size, (KFR_DFT_PACK_FORMAT)(int)fmt);
        amd::check();
        return n;
    }
    progressive progressive_start(T* out, const complex<T>* in, u8* temp) const { return start(true, out, in, temp); }
    progressive progressive_start(complex<T>* out, const T* in, u8* temp) const { return start(false, out, in, temp); }
    bool progressive_step(progressive& prog) const
    {
        if (!prog.state)
            throw logic_error("progressive execution was not started");
        const int r = kfr_amd_progressive_step(prog.state->h);
        amd::check(r >= 0);
        ++prog.step;
        return r == 1;
    }

    typename api::real_plan* native() const { return h.get(); }

protected:
    friend struct dft_cache_impl_t<T>;
    struct adopt
    {
    };
    dft_plan_real(adopt, typename api::real_plan* p, size_t n, dft_pack_format f) : size(n), fmt(f), h(p)
    {
        amd::check(p != nullptr);
        temp_size = api::rtemp(p);
    }
    progressive start(bool inverse, void* out, const void* in, u8* temp) const
    {
        if (!is_initialized())
            throw logic_error("dft_plan_real is not initialized");
        if (temp_size != 0 && temp == nullptr)
            throw logic_error("Temporary buffer must be provided for progressive execution");
        progressive r;
        r.inverse = inverse;
        r.temp    = temp;
        KFR_AMD_PROGRESSIVE* p = api::rprog(h.get(), inverse, out, in);
        amd::check(p != nullptr);
        r.state = std::make_shared<amd::progressive_handle>(p);
        return r;
    }
    amd::handle<typename api::real_plan, &api::rdel> h;
};

/// DCT-II, unscaled; the inverse is DCT-III with x[0] halved (include/kfr/dft/fft.hpp:927-978)
template <typename T>
struct dct_plan
{
    using api = amd::capi<T>;
    size_t size = 0;
    size_t temp_size = 0;
    dct_plan(const dct_plan&)            = delete;
    dct_plan& operator=(const dct_plan&) = delete;
    dct_plan(dct_plan&&)                 = default;
    dct_plan& operator=(dct_plan&&)      = default;
    dct_plan(size_t size) : size(size), h(api::dcreate(size))
    {
        amd::check(h.h != nullptr);
        temp_size = api::dtemp(h.h);
    }
    void dump() const { api::ddump(h.get()); }
    void execute(T* out, const T* in, u8* temp, bool inverse = false) const
    {
        api::dexec(h.get(), out, in, temp, inverse);
        amd::check();
    }
    template <univector_tag Tag1, univector_tag Tag2, univector_tag Tag3>
    void execute(univector<T, Tag1>& out, const univector<T, Tag2>& in, univector<u8, Tag3>& temp,
                 bool inverse = false) const
    {
        execute(out.data(), in.data(), temp.data(), inverse);
    }

private:
    amd::handle<typename api::dct_plan, &api::ddel> h;
};

/// Multidimensional complex DFT over a row-major array (include/kfr/dft/fft.hpp:569-691)
template <typename T>
struct dft_plan_md
{
    using api = amd::capi<T>;
    std::vector<size_t> size;
    size_t temp_size = 0;
    dft_plan_md(const dft_plan_md&)            = delete;
    dft_plan_md& operator=(const dft_plan_md&) = delete;
    dft_plan_md(dft_plan_md&&)                 = default;
    dft_plan_md& operator=(dft_plan_md&&)      = default;
    explicit dft_plan_md(std::vector<size_t> shape) : size(std::move(shape))
    {
        std::vector<unsigned> s(size.begin(), size.end());
        h = decltype(h)(api::create_md(s.size(), s.data()));
        amd::check(h.h != nullptr);
        temp_size = api::temp(h.h);
    }
    dft_plan_md(std::initializer_list<size_t> shape) : dft_plan_md(std::vector<size_t>(shape)) {}
    bool is_initialized() const
    {
        size_t p = size.empty() ? 0 : 1;
        for (size_t v : size)
            p *= v;
        return p != 0;
    }
    void dump() const { api::dump(h.get()); }
    void execute(complex<T>* out, const complex<T>* in, u8* temp, bool inverse = false) const
    {
        api::exec(h.get(), out, in, temp, inverse);
        amd::check();
    }
    template <bool inverse = false>
    void execute(complex<T>* out, const complex<T>* in, u8* temp, cbool_t<inverse>) const
    {
        execute(out, in, temp, inverse);
    }

private:
    amd::handle<typename api::plan, &api::del> h;
};

/// Multidimensional real DFT: real input of `size`, complex output with the last axis n/2+1 (CCs)
/// (include/kfr/dft/fft.hpp:693-924)
template <typename T>
struct dft_plan_md_real
{
    using api = amd::capi<T>;
    std::vector<size_t> size;
    size_t temp_size = 0;
    bool real_out_is_enough = false;
    dft_plan_md_real(const dft_plan_md_real&)            = delete;
    dft_plan_md_real& operator=(const dft_plan_md_real&) = delete;
    dft_plan_md_real(dft_plan_md_real&&)                 = default;
    dft_plan_md_real& operator=(dft_plan_md_real&&)      = default;
    explicit dft_plan_md_real(std::vector<size_t> shape, bool real_out_is_enough = false)
        : size(std::move(shape)), real_out_is_enough(real_out_is_enough)
    {
        std::vector<unsigned> s(size.begin(), size.end());
        h = decltype(h)(api::rcreate_md(s.size(), s.data(), real_out_is_enough));
        amd::check(h.h != nullptr);
        temp_size = api::rtemp(h.h);
    }
    std::vector<size_t> complex_size() const
    {
        std::vector<size_t> c = size;
        if (!c.empty())
            c.back() = c.back() / 2 + 1;
        return c;
    }
    size_t complex_size_flat() const
    {
        size_t p = 1;
        for (size_t v : complex_size())
            p *= v;
        return p;
    }
    void dump() const { api::rdump(h.get()); }
    void execute(complex<T>* out, const T* in, u8* temp, cdirect_t = {}) const
    {
        api::rfwd(h.get(), out, in, temp);
        amd::check();
    }
    void execute(T* out, const complex<T>* in, u8* temp, cinvert_t = {}) const
    {
        api::rinv(h.get(), out, in, temp);
        amd::check();
    }

private:
    amd::handle<typename api::real_plan, &api::rdel> h;
};

// ---- convolution (include/kfr/dft/convolution.hpp:53-76)
namespace amd
{
template <typename T>
struct conv_types
{
    using sub                   = T;
    static constexpr int dtype  = std::is_same_v<T, float> ? 0 : 1;
};
template <typename T>
struct conv_types<complex<T>>
{
    using sub                   = T;
    static constexpr int dtype  = std::is_same_v<T, float> ? 2 : 3;
};
template <typename T>
univector<T> convolve(int mode, const T* a, size_t na, const T* b, size_t nb)
{
    const long n = kfr_amd_convolve(conv_types<T>::dtype, mode, a, na, b, nb, nullptr);
    check(n >= 0);
    univector<T> out((size_t)n);
    check(kfr_amd_convolve(conv_types<T>::dtype, mode, a, na, b, nb, out.data()) >= 0);
    return out;
}
} // namespace amd

/// Convolution; for complex T the result keeps only the real parts, as the reference's does
template <typename T1, typename T2, univector_tag Tag1, univector_tag Tag2>
    requires(std::is_same_v<std::remove_const_t<T1>, std::remove_const_t<T2>>)
univector<std::remove_const_t<T1>> convolve(const univector<T1, Tag1>& src1, const univector<T2, Tag2>& src2)
{
    return amd::convolve<std::remove_const_t<T1>>(0, src1.data(), src1.size(), src2.data(), src2.size());
}
/// Correlation: convolution with src2 reversed
template <typename T1, typename T2, univector_tag Tag1, univector_tag Tag2>
    requires(std::is_same_v<std::remove_const_t<T1>, std::remove_const_t<T2>>)
univector<std::remove_const_t<T1>> correlate(const univector<T1, Tag1>& src1, const univector<T2, Tag2>& src2)
{
    return amd::convolve<std::remove_const_t<T1>>(1, src1.data(), src1.size(), src2.data(), src2.size());
}
/// Auto-correlation: the second half of correlate(src, src)
template <typename T, univector_tag Tag1>
univector<std::remove_const_t<T>> autocorrelate(const univector<T, Tag1>& src)
{
    return amd::convolve<std::remove_const_t<T>>(2, src.data(), src.size(), src.data(), src.size());
}

/// Streaming FFT convolution filter (include/kfr/dft/convolution.hpp:99-148), real or complex T
template <typename T>
class convolve_filter
{
    using ST   = typename amd::conv_types<T>::sub;
    using api  = amd::capi<ST>;
    static constexpr bool real_fft = std::is_same_v<T, ST>;

public:
    explicit convolve_filter(size_t size, size_t block_size = 1024)
        : convolve_filter(univector<T>(size, T(0)).slice(), block_size)
    {
    }
    explicit convolve_filter(const univector_ref<const T>& data, size_t block_size = 1024) : bs(block_size)
    {
        if constexpr (real_fft)
            rf = decltype(rf)(api::fcreate(data.data(), data.size(), block_size));
        else
            cf = decltype(cf)(api::cfcreate(data.data(), data.size(), block_size));
        amd::check(real_fft ? rf.h != nullptr : cf.h != nullptr);
        size_t b = 1;
        while (b < (block_size ? block_size : 1024))
            b <<= 1;
        bs = b;
    }
    template <univector_tag Tag>
    explicit convolve_filter(const univector<T, Tag>& data, size_t block_size = 1024)
        : convolve_filter(univector_ref<const T>(data.data(), data.size()), block_size)
    {
    }
    void set_data(const univector_ref<const T>& data)
    {
        if constexpr (real_fft)
            api::fset(rf.get(), data.data(), data.size());
        else
            api::cfset(cf.get(), data.data(), data.size());
        amd::check();
    }
    void reset()
    {
        if constexpr (real_fft)
            api::freset(rf.get());
        else
            api::cfreset(cf.get());
        amd::check();
    }
    /// Apply filter to multiples of returned block size for optimal processing efficiency.
    size_t input_block_size() const { return bs; }

    /// filter `size` samples of src into dest (in place allowed); the state carries across calls
    void apply(T* dest, const T* src, size_t size)
    {
        if constexpr (real_fft)
            api::fproc(rf.get(), dest, src, size);
        else
            api::cfproc(cf.get(), dest, src, size);
        amd::check();
    }
    template <univector_tag Tag>
    void apply(univector<T, Tag>& dest)
    {
        apply(dest.data(), dest.data(), dest.size());
    }
    template <univector_tag Tag1, univector_tag Tag2>
    void apply(univector<T, Tag1>& dest, const univector<T, Tag2>& src)
    {
        apply(dest.data(), src.data(), std::min(dest.size(), src.size()));
    }

private:
    size_t bs;
    amd::handle<typename api::filter, &api::fdel> rf;
    amd::handle<typename api::cfilter, &api::cfdel> cf;
};

// ---- plan cache (include/kfr/dft/cache.hpp:44-170)
template <typename T>
using dft_plan_ptr = std::shared_ptr<const dft_plan<T>>;
template <typename T>
using dft_plan_real_ptr = std::shared_ptr<const dft_plan_real<T>>;

template <typename T>
struct dft_cache_impl_t
{
    static dft_plan_ptr<T> get(size_t size)
    {
        return std::shared_ptr<const dft_plan<T>>(
            new dft_plan<T>(typename dft_plan<T>::adopt{}, amd::capi<T>::cached(size)));
    }
    static dft_plan_real_ptr<T> getreal(size_t size, dft_pack_format fmt)
    {
        return std::shared_ptr<const dft_plan_real<T>>(new dft_plan_real<T>(
            typename dft_plan_real<T>::adopt{}, amd::capi<T>::rcached(size, (int)fmt), size, fmt));
    }
};

/// The process-wide plan cache: plans are shared with every caller of the same key
/// (the library's cache, kfr_amd_cache_*)
struct dft_cache
{
    static dft_cache& instance()
    {
        static dft_cache c;
        return c;
    }
    dft_plan_ptr<f32> get(ctype_t<f32>, size_t size) { return dft_cache_impl_t<f32>::get(size); }
    dft_plan_ptr<f64> get(ctype_t<f64>, size_t size) { return dft_cache_impl_t<f64>::get(size); }
    dft_plan_real_ptr<f32> getreal(ctype_t<f32>, size_t size)
    {
        return dft_cache_impl_t<f32>::getreal(size, dft_pack_format::CCs);
    }
    dft_plan_real_ptr<f64> getreal(ctype_t<f64>, size_t size)
    {
        return dft_cache_impl_t<f64>::getreal(size, dft_pack_format::CCs);
    }
    void clear()
    {
        kfr_amd_cache_clear();
        amd::check();
    }
    size_t size() const { return kfr_amd_cache_size(); }
};

/// Direct DFT through the cached plan
template <typename T, univector_tag Tag>
univector<complex<T>> dft(const univector<complex<T>, Tag>& input)
{
    dft_plan_ptr<T> d = dft_cache::instance().get(ctype_t<T>(), input.size());
    univector<complex<T>> output(input.size(), complex<T>(std::numeric_limits<T>::quiet_NaN()));
    univector<u8> temp(d->temp_size);
    d->execute(output, input, temp);
    return output;
}
/// Inverse DFT through the cached plan (unscaled)
template <typename T, univector_tag Tag>
univector<complex<T>> idft(const univector<complex<T>, Tag>& input)
{
    dft_plan_ptr<T> d = dft_cache::instance().get(ctype_t<T>(), input.size());
    univector<complex<T>> output(input.size(), complex<T>(std::numeric_limits<T>::quiet_NaN()));
    univector<u8> temp(d->temp_size);
    d->execute(output, input, temp, ctrue);
    return output;
}
/// Real direct DFT through the cached plan (CCs, size/2+1 values)
template <typename T, univector_tag Tag>
univector<complex<T>> realdft(const univector<T, Tag>& input)
{
    dft_plan_real_ptr<T> d = dft_cache::instance().getreal(ctype_t<T>(), input.size());
    univector<complex<T>> output(input.size() / 2 + 1, complex<T>(std::numeric_limits<T>::quiet_NaN()));
    univector<u8> temp(d->temp_size);
    d->execute(output, input, temp);
    return output;
}
/// Real inverse DFT through the cached plan of (size-1)*2 points (unscaled)
template <typename T, univector_tag Tag>
univector<T> irealdft(const univector<complex<T>, Tag>& input)
{
    dft_plan_real_ptr<T> d = dft_cache::instance().getreal(ctype_t<T>(), (input.size() - 1) * 2);
    univector<T> output((input.size() - 1) * 2, std::numeric_limits<T>::quiet_NaN());
    univector<u8> temp(d->temp_size);
    d->execute(output, input, temp);
    return output;
}

} // namespace kfr
