// KFR plan descriptor: the stage list a reference dft_plan / dft_plan_real / dct_plan builds for a
// given size and element type, restated on the host so that dump(), temp_size and the reorder
// permutations of a GPU plan are identical to the reference's. The GPU schedule is lowered from the
// size separately (gpu_plan.hpp); this descriptor is the parity record, not the execution recipe.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace kfr_amd
{

enum class dtype_t : int
{
    f32 = 0,
    f64 = 1
};

inline size_t complex_bytes(dtype_t t) { return t == dtype_t::f32 ? 8 : 16; }

// One reference dft_stage (include/kfr/dft/fft.hpp:58-81): the fields dump() prints.
struct stage_desc
{
    std::string name; // template name without the "(arch)" suffix
    size_t radix      = 0;
    size_t stage_size = 0;
    size_t data_size  = 0;
    size_t temp_size  = 0;
    size_t repeats    = 1;
    size_t out_offset = 0;
    size_t blocks     = 0;
    bool recursion    = false;
    bool can_inplace  = true;
    std::vector<stage_desc> nested; // dft_special_stage_impl prints its two inner stages
};

enum class reorder_kind : int
{
    none      = 0,
    bitrev    = 1, // fft_reorder_stage_impl (src/dft/fft-impl.hpp:982-1001)
    digitrev  = 2  // dft_reorder_stage_impl (src/dft/dft-impl.hpp:428-494)
};

struct plan_desc
{
    dtype_t dtype   = dtype_t::f32;
    size_t size     = 0; // complex size of the underlying dft_plan (N/2 for real plans)
    size_t temp_size = 0;
    size_t data_size = 0;
    bool arblen     = false;
    std::vector<stage_desc> stages; // all_stages
    std::vector<int> order[2];      // stages[0] (forward) and stages[1] (inverse) as indexes
    std::vector<int> radices;       // init_dft factorization (ascending), empty for power-of-2 sizes
    reorder_kind reorder = reorder_kind::none;
    uint64_t disposition_inplace[2]    = { 0, 0 };
    uint64_t disposition_outofplace[2] = { 0, 0 };
};

// dft_plan<T>(size, dft_order::normal, progressive_optimized)  (include/kfr/dft/fft.hpp:242-246)
plan_desc describe_dft(dtype_t t, size_t size, bool progressive_optimized = false);
// dft_plan_real<T>(size, fmt)  (include/kfr/dft/fft.hpp:479-487); fmt 0 = Perm, 1 = CCs
plan_desc describe_dft_real(dtype_t t, size_t size, int fmt);
// dct_plan<T>(size)  (include/kfr/dft/fft.hpp:930)
plan_desc describe_dct(dtype_t t, size_t size);

// Text in the layout of dft_plan::dump() (src/dft/dft.cpp:35-42 and fft.hpp:77-81), one line per stage,
// with `arch` in the place of the reference's ISA name.
std::string dump_text(const plan_desc& d, const char* arch);

// The permutation the plan's reorder stage applies: out[i] = in[perm[i]]. Empty when the plan has no
// reorder stage.
std::vector<int64_t> reorder_permutation(const plan_desc& d);

// Helpers shared with the GPU lowering.
inline bool is_pow2(size_t n) { return n != 0 && (n & (n - 1)) == 0; }
inline size_t next_pow2(size_t n)
{
    size_t p = 1;
    while (p < n)
        p <<= 1;
    return p;
}
inline int ilog2(size_t n)
{
    int l = 0;
    while ((size_t(1) << (l + 1)) <= n)
        ++l;
    return l;
}

} // namespace kfr_amd
