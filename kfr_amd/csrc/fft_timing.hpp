// Per-launch timing of the library's transform kernels with HIP events on the launching stream.
// Off by default; while a session is open, each instrumented launch is bracketed by two events and
// recorded with its kernel name (the template instance, spelt as profilers print it) and its
// algorithmic bytes (one read of the data it transforms and one write of its result; tables
// excluded). Reading the session synchronises its events.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

namespace kfr_amd
{
namespace timing
{

struct record
{
    std::string name;
    long long bytes;
    hipEvent_t start, stop;
};

struct state
{
    std::mutex m;
    std::atomic<bool> on{ false };
    std::vector<record> recs;
    std::vector<hipEvent_t> pool;
};

inline state& get()
{
    static state s;
    return s;
}

inline bool enabled() { return get().on.load(std::memory_order_relaxed); }

inline hipEvent_t take_event(state& st)
{
    if (!st.pool.empty())
    {
        hipEvent_t e = st.pool.back();
        st.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess)
        return nullptr;
    return e;
}

// brackets one kernel launch: construct right before it, the destructor records the end event
class span
{
public:
    template <typename NameFn>
    span(NameFn&& name, long long bytes, hipStream_t s) : s_(s)
    {
        if (!enabled())
            return;
        state& st = get();
        std::lock_guard<std::mutex> g(st.m);
        hipEvent_t a = take_event(st), b = take_event(st);
        if (!a || !b)
            return;
        if (hipEventRecord(a, s) != hipSuccess)
            return;
        rec_ = (long long)st.recs.size();
        st.recs.push_back(record{ name(), bytes, a, b });
    }
    ~span()
    {
        if (rec_ < 0)
            return;
        state& st = get();
        std::lock_guard<std::mutex> g(st.m);
        if (rec_ < (long long)st.recs.size())
            (void)hipEventRecord(st.recs[rec_].stop, s_);
    }
    span(const span&)            = delete;
    span& operator=(const span&) = delete;

private:
    hipStream_t s_;
    long long rec_ = -1;
};

template <typename T>
inline const char* type_name()
{
    return sizeof(T) == 4 ? "float" : "double";
}

inline void append(std::string& s, const char* v) { s += v; }
inline void append(std::string& s, bool v) { s += v ? "true" : "false"; }
inline void append(std::string& s, int v) { s += std::to_string(v); }

// "k<a, b, c>", the spelling of a template instance in profiler output
template <typename... A>
std::string kernel_name(const char* k, A... a)
{
    std::string s = k;
    s += '<';
    bool first = true;
    ((s += first ? "" : ", ", first = false, append(s, a)), ...);
    s += '>';
    return s;
}

} // namespace timing
} // namespace kfr_amd
