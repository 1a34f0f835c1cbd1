// CPU emulation of the subset of the HIP runtime and device language that the kfr_amd engine uses, so
// that its kernels can run on the host under AddressSanitizer (every LDS and global access checked
// against the exact allocation) and ThreadSanitizer (a missing barrier shows as a data race).
//
// A kernel launch runs its workgroups one after another; each workgroup runs as blockDim host
// threads that meet at __syncthreads (and at the LDS barrier builtin) on one std::barrier, and a
// thread that returns drops out of the barrier, as an ended wave does on the GPU. The dynamic LDS of a
// workgroup is one heap block of exactly the launch's byte count. Device memory is host memory.
// Wave-level operations (__shfl_xor, v_permlane32_swap, the 32x32x16 bf16 MFMA) exchange operands
// through a per-wave buffer between two wave barriers; the MFMA follows the lane layout of
// v_mfma_f32_32x32x16_bf16 (lane l, slot s: A[l%32][8(l/32)+s], B[8(l/32)+s][l%32]; result slot v:
// D[8(v/4)+4(l/32)+v%4][l%32]) with the products summed in FP32.
#pragma once

#include <algorithm>
#include <barrier>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <climits>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__
#define __align__(n)

struct dim3
{
    uint32_t x, y, z;
    constexpr dim3(uint32_t a = 1, uint32_t b = 1, uint32_t c = 1) : x(a), y(b), z(c) {}
};

inline thread_local dim3 threadIdx;
inline thread_local dim3 blockIdx;
inline thread_local dim3 blockDim;
inline thread_local dim3 gridDim;

namespace emu
{
inline thread_local std::barrier<>* cur_barrier = nullptr;
inline thread_local unsigned char* cur_lds      = nullptr;
inline void sync()
{
    if (cur_barrier)
        cur_barrier->arrive_and_wait();
}
[[noreturn]] inline void unsupported(const char* what)
{
    std::fprintf(stderr, "cpu emulation: %s is not emulated\n", what);
    std::abort();
}
} // namespace emu

// The dynamic LDS array of the running workgroup. Kernels declare it at block scope as
// `extern __shared__ ... smem_raw[]`, which here becomes a block-scope declaration of the function
// emu_smem() returning a pointer to the array; such a declaration names a function of the enclosing
// namespace, so every namespace that holds kernels gets its own definition.
#define EMU_SMEM_IN(ns)                                                                                    \
    namespace ns                                                                                           \
    {                                                                                                      \
    inline unsigned char (*emu_smem())[] { return reinterpret_cast<unsigned char (*)[]>(emu::cur_lds); } \
    }
EMU_SMEM_IN(kfr_amd)
EMU_SMEM_IN(kfr_amd::p2)
EMU_SMEM_IN(kfr_amd::mx)
EMU_SMEM_IN(kfr_amd::bt)
#define smem_raw (*emu_smem())

#define __syncthreads() ::emu::sync()
#define __builtin_amdgcn_s_barrier() ::emu::sync()
#define __builtin_amdgcn_fence(...) ((void)0)
#define __builtin_amdgcn_sched_barrier(...) ((void)0)
typedef unsigned emu_u32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 emu_bf16x8 __attribute__((ext_vector_type(8)));
typedef float emu_f32x16 __attribute__((ext_vector_type(16)));

namespace emu
{
// the lanes of one wave: a barrier of its live lanes and a slot per lane for exchanged operands
struct wave_t
{
    std::barrier<> bar;
    alignas(64) unsigned char slot[64][128];
    explicit wave_t(std::ptrdiff_t lanes) : bar(lanes) {}
};
inline thread_local wave_t* cur_wave = nullptr;
inline int lane() { return (int)(threadIdx.x % 64); }

// every lane publishes `mine`, then reads lane src's
template <typename V>
V wave_read(const V& mine, int src)
{
    static_assert(sizeof(V) <= 128, "operand too large");
    std::memcpy(cur_wave->slot[lane()], &mine, sizeof(V));
    cur_wave->bar.arrive_and_wait();
    V r;
    std::memcpy(&r, cur_wave->slot[src], sizeof(V));
    cur_wave->bar.arrive_and_wait();
    return r;
}

inline emu_u32x2 permlane32_swap(unsigned v0, unsigned v1)
{
    const int l = lane();
    struct pair
    {
        unsigned a, b;
    };
    const pair other = wave_read(pair{ v0, v1 }, l ^ 32);
    emu_u32x2 r;
    r[0] = l >= 32 ? other.b : v0; // the upper half of v0 takes the lower half of v1
    r[1] = l < 32 ? other.a : v1;  // the lower half of v1 takes the upper half of v0
    return r;
}

inline emu_f32x16 mfma_32x32x16_bf16(emu_bf16x8 a, emu_bf16x8 b, emu_f32x16 c)
{
    struct ops
    {
        float a[8], b[8];
    };
    ops mine;
    for (int s = 0; s < 8; ++s)
    {
        mine.a[s] = (float)a[s];
        mine.b[s] = (float)b[s];
    }
    std::memcpy(cur_wave->slot[lane()], &mine, sizeof(ops));
    cur_wave->bar.arrive_and_wait();
    float A[32][16], B[16][32];
    for (int l = 0; l < 64; ++l)
    {
        ops o;
        std::memcpy(&o, cur_wave->slot[l], sizeof(ops));
        for (int s = 0; s < 8; ++s)
        {
            A[l % 32][8 * (l / 32) + s] = o.a[s];
            B[8 * (l / 32) + s][l % 32] = o.b[s];
        }
    }
    cur_wave->bar.arrive_and_wait();
    const int l = lane();
    emu_f32x16 d;
    for (int v = 0; v < 16; ++v)
    {
        const int i = 8 * (v / 4) + 4 * (l / 32) + v % 4, j = l % 32;
        float acc   = c[v];
        for (int k = 0; k < 16; ++k)
            acc += A[i][k] * B[k][j];
        d[v] = acc;
    }
    return d;
}
} // namespace emu

#define __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, B, C, ...) ::emu::mfma_32x32x16_bf16(A, B, C)
#define __builtin_amdgcn_permlane32_swap(V0, V1, ...) ::emu::permlane32_swap(V0, V1)

inline int __lane_id() { return emu::lane(); }
template <typename V>
inline V __shfl_xor(V v, int mask)
{
    return emu::wave_read(v, emu::lane() ^ mask);
}
inline unsigned __float_as_uint(float f)
{
    unsigned u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float __uint_as_float(unsigned u)
{
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
inline uint32_t __umulhi(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }
inline void sincospi(double x, double* s, double* c)
{
    *s = std::sin(M_PI * x);
    *c = std::cos(M_PI * x);
}

struct alignas(8) float2
{
    float x, y;
};
struct alignas(16) float4
{
    float x, y, z, w;
};
struct alignas(16) double2
{
    double x, y;
};

// ---- runtime

enum hipError_t
{
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2
};
inline const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "emulated error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }

struct emu_stream_t;
struct emu_event_t;
struct emu_graph_t;
typedef emu_stream_t* hipStream_t;
typedef emu_event_t* hipEvent_t;
typedef emu_graph_t* hipGraph_t;
typedef emu_graph_t* hipGraphExec_t;

enum hipMemcpyKind
{
    hipMemcpyHostToHost,
    hipMemcpyHostToDevice,
    hipMemcpyDeviceToHost,
    hipMemcpyDeviceToDevice,
    hipMemcpyDefault
};
enum hipStreamCaptureStatus
{
    hipStreamCaptureStatusNone,
    hipStreamCaptureStatusActive
};
enum hipFuncAttribute
{
    hipFuncAttributeMaxDynamicSharedMemorySize
};
enum hipDeviceAttribute_t
{
    hipDeviceAttributeMultiprocessorCount
};
constexpr unsigned hipStreamNonBlocking   = 1;
constexpr unsigned hipEventDisableTiming  = 2;

inline hipError_t hipMalloc(void** p, size_t n)
{
    *p = std::malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
template <typename T>
inline hipError_t hipMalloc(T** p, size_t n)
{
    return hipMalloc(reinterpret_cast<void**>(p), n);
}
inline hipError_t hipFree(void* p)
{
    std::free(p);
    return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind)
{
    std::memmove(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t = nullptr)
{
    std::memmove(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t = nullptr)
{
    std::memset(d, v, n);
    return hipSuccess;
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned)
{
    *s = nullptr;
    return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus* c)
{
    *c = hipStreamCaptureStatusNone;
    return hipSuccess;
}
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e)
{
    *e = nullptr;
    return hipErrorInvalidValue; // no timing in the emulation
}
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t = nullptr) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t)
{
    *ms = 0.f;
    return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipFuncSetAttribute(const void*, hipFuncAttribute, int) { return hipSuccess; }
inline hipError_t hipGetDevice(int* d)
{
    *d = 0;
    return hipSuccess;
}
inline hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int)
{
    *v = 256;
    return hipSuccess;
}

namespace emu
{
// largest dynamic LDS a workgroup may use (gfx950: 160 KiB)
constexpr size_t max_lds = 160 * 1024;

// Persistent host threads, one per thread of the largest workgroup launched so far: the launcher
// publishes one workgroup at a time and waits until every one of its threads has returned.
class pool
{
public:
    static pool& get()
    {
        static pool p;
        return p;
    }
    template <typename Body>
    void run_block(unsigned nt, Body&& body)
    {
        grow(nt);
        std::barrier<> bar((std::ptrdiff_t)nt);
        {
            std::lock_guard<std::mutex> l(mu_);
            job_     = [&](unsigned t)
            {
                cur_barrier = &bar;
                body(t);
                bar.arrive_and_drop();
            };
            nt_      = nt;
            pending_ = nt;
            ++gen_;
        }
        cv_.notify_all();
        std::unique_lock<std::mutex> l(mu_);
        done_cv_.wait(l, [&] { return pending_ == 0; });
    }
    ~pool()
    {
        {
            std::lock_guard<std::mutex> l(mu_);
            stop_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (auto& t : th_)
            t.join();
    }

private:
    // called before the next job is published: a new worker starts from the current generation, so it
    // takes that job
    void grow(unsigned nt)
    {
        while (th_.size() < nt)
        {
            const unsigned id = (unsigned)th_.size();
            unsigned long long g;
            {
                std::lock_guard<std::mutex> l(mu_);
                g = gen_;
            }
            th_.emplace_back([this, id, g] { worker(id, g); });
        }
    }
    void worker(unsigned id, unsigned long long seen)
    {
        for (;;)
        {
            std::function<void(unsigned)> job;
            {
                std::unique_lock<std::mutex> l(mu_);
                cv_.wait(l, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_)
                    return;
                if (id >= nt_)
                    continue;
                job = job_;
            }
            job(id);
            std::lock_guard<std::mutex> l(mu_);
            if (--pending_ == 0)
                done_cv_.notify_one();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::vector<std::thread> th_;
    std::function<void(unsigned)> job_;
    unsigned nt_ = 0, pending_ = 0;
    unsigned long long gen_ = 0;
    bool stop_ = false;
};

template <typename K, typename... A>
void launch(K kernel, dim3 grid, dim3 block, size_t shmem, hipStream_t, A... args)
{
    const unsigned nt = block.x * block.y * block.z;
    if (nt == 0 || nt > 1024)
    {
        std::fprintf(stderr, "cpu emulation: workgroup of %u threads\n", nt);
        std::abort();
    }
    if (shmem > max_lds)
    {
        std::fprintf(stderr, "cpu emulation: %zu bytes of LDS requested\n", shmem);
        std::abort();
    }
    for (uint32_t bz = 0; bz < grid.z; ++bz)
        for (uint32_t by = 0; by < grid.y; ++by)
            for (uint32_t bx = 0; bx < grid.x; ++bx)
            {
                std::unique_ptr<unsigned char[]> lds(shmem ? new unsigned char[shmem] : nullptr);
                if (shmem)
                    std::memset(lds.get(), 0xA5, shmem); // stale contents, as on the GPU
                unsigned char* l = lds.get();
                std::vector<std::unique_ptr<wave_t>> waves;
                for (unsigned w = 0; w * 64 < nt; ++w)
                    waves.emplace_back(new wave_t((std::ptrdiff_t)std::min(64u, nt - 64 * w)));
                pool::get().run_block(nt,
                                      [&](unsigned t)
                                      {
                                          threadIdx = dim3(t % block.x, (t / block.x) % block.y,
                                                           t / (block.x * block.y));
                                          blockIdx  = dim3(bx, by, bz);
                                          blockDim  = block;
                                          gridDim   = grid;
                                          cur_lds   = l;
                                          cur_wave  = waves[t / 64].get();
                                          kernel(args...);
                                          cur_wave->bar.arrive_and_drop();
                                      });
            }
}
} // namespace emu

#define hipLaunchKernelGGL(K, G, B, S, ST, ...) ::emu::launch(K, dim3(G), dim3(B), (size_t)(S), ST, ##__VA_ARGS__)
