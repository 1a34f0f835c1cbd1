// Device scratch of the entry point that allocates nothing (kfr_amd_dft_execute_batch_noalloc_f32):
// one file-scope device array, which every GPU the library runs on holds until the process ends. Its
// device address differs per GPU and is looked up once per device, with that device current.
#include "fft_engine.hpp"

#include <hip/hip_runtime.h>

#include <mutex>

namespace kfr_amd
{

static __device__ __align__(256) unsigned char scratch[noalloc_scratch_bytes];

workspace noalloc_workspace()
{
    constexpr int max_devices = 64;
    static std::mutex mu;
    static void* ptr[max_devices] = {};
    int dev                       = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= max_devices)
        throw std::runtime_error("device index out of range for the allocation-free workspace");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev])
    {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
        ptr[dev] = p;
    }
    return workspace{ ptr[dev], noalloc_scratch_bytes };
}

} // namespace kfr_amd
