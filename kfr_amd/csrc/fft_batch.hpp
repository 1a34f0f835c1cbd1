// Uneven batches: many complex transforms of different lengths and directions in one launch.
#pragma once

#include "fft_engine.hpp"

#include <memory>
#include <string>

namespace kfr_amd
{

// Transform i has sizes[i] points and direction inverse[i]; its data starts at the prefix sum of the
// sizes before it (in both the input and the output buffer). Every length must fit one workgroup
// (up to 8192 points for f32, 4096 for f64, Bluestein lengths up to half that).
template <typename T>
class gpu_batch
{
public:
    // autotune: time candidate schedules of every distinct length on the device (pass order, LDS
    // padding, transforms per work item; cached per process) instead of the planner's order, one
    // padding slot per 16 values and the occupancy heuristic
    gpu_batch(size_t count, const size_t* sizes, const int* inverse, bool autotune);
    ~gpu_batch();
    size_t count() const;
    size_t total() const; // complex values in the input (and output)
    // one graph launch (queue reset + persistent kernel) on stream s; in and out are device buffers,
    // in == out allowed. Calls on one plan must not overlap (they share the work-queue counter).
    void execute(void* out, const void* in, hipStream_t s);
    std::string describe() const;

private:
    struct impl;
    std::unique_ptr<impl> p;
};

extern template class gpu_batch<float>;
extern template class gpu_batch<double>;

} // namespace kfr_amd
