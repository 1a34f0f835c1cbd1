/*
 * C ABI of the MI355X (gfx950) DFT engine.
 *
 * Part 1 is a drop-in for the reference's C API (include/kfr/capi.h): the same function names,
 * argument meanings, opaque plan types, pack-format enum and error channel, so a program written
 * against the reference links against libkfr_amd.so unchanged. Each declaration names the
 * reference declaration it replaces. Semantics kept from the reference:
 *   - plans are created by *_create_* and freed by *_delete_plan_*; NULL is returned on failure and
 *     kfr_last_error() describes why; every successful call clears the error (src/capi/capi.cpp:36-91);
 *   - execute functions take host pointers (device pointers are also accepted), run synchronously,
 *     apply no scaling, and allow out == in; `temp` may be NULL (the GPU keeps its own scratch);
 *   - *_get_temp_size returns the same byte count the reference plan reports.
 *
 * Part 2 holds extensions the reference does not have: batched transforms on device pointers with
 * a HIP stream, FFT convolution with its adjoint, plan descriptors and the GPU schedule.
 */
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C"
{
#endif

#define KFR_AMD_API __attribute__((visibility("default")))

/* include/kfr/capi.h:77-96 */
#if defined(__cplusplus)
typedef bool kfr_bool;
#else
typedef _Bool kfr_bool;
#endif

/* include/kfr/capi.h:98-111 */
enum
{
    KFR_ARCH_X86    = 0,
    KFR_ARCH_SSE2   = 1,
    KFR_ARCH_SSE3   = 2,
    KFR_ARCH_SSSE3  = 3,
    KFR_ARCH_SSE41  = 4,
    KFR_ARCH_SSE42  = 5,
    KFR_ARCH_AVX    = 6,
    KFR_ARCH_AVX2   = 7,
    KFR_ARCH_AVX512 = 8,
    KFR_ARCH_GFX950 = 100 /* this library's only target */
};

#define KFR_HEADERS_VERSION 70000 /* include/kfr/capi.h:113 */

/* include/kfr/capi.h:131-142: complex values are interleaved (re, im) pairs */
typedef float kfr_f32;
typedef double kfr_f64;
typedef float kfr_c32;
typedef double kfr_c64;
#define KFR_COMPLEX_SIZE_MULTIPLIER 2

/* include/kfr/capi.h:145-164 */
typedef struct KFR_DFT_PLAN_F32 { int opaque; } KFR_DFT_PLAN_F32;
typedef struct KFR_DFT_PLAN_F64 { int opaque; } KFR_DFT_PLAN_F64;
typedef struct KFR_DFT_REAL_PLAN_F32 { int opaque; } KFR_DFT_REAL_PLAN_F32;
typedef struct KFR_DFT_REAL_PLAN_F64 { int opaque; } KFR_DFT_REAL_PLAN_F64;
typedef struct KFR_DCT_PLAN_F32 { int opaque; } KFR_DCT_PLAN_F32;
typedef struct KFR_DCT_PLAN_F64 { int opaque; } KFR_DCT_PLAN_F64;
typedef struct KFR_FILTER_F32 { int opaque; } KFR_FILTER_F32;
typedef struct KFR_FILTER_F64 { int opaque; } KFR_FILTER_F64;
typedef struct KFR_FILTER_C32 { int opaque; } KFR_FILTER_C32;
typedef struct KFR_FILTER_C64 { int opaque; } KFR_FILTER_C64;

#define KFR_DEFAULT_ALIGNMENT 64 /* include/kfr/capi.h:167 */

/* include/kfr/capi.h:196-200 */
typedef enum KFR_DFT_PACK_FORMAT
{
    Perm = 0,
    CCs  = 1
} KFR_DFT_PACK_FORMAT;

/* ============================== Part 1: drop-in ============================== */

/* library information: include/kfr/capi.h:116-128 */
KFR_AMD_API const char* kfr_version_string(void);
KFR_AMD_API uint32_t kfr_version(void);
KFR_AMD_API const char* kfr_enabled_archs(void);
KFR_AMD_API int kfr_current_arch(void);
KFR_AMD_API const char* kfr_last_error(void);

/* host memory: include/kfr/capi.h:170-176 */
KFR_AMD_API void* kfr_allocate(size_t size);
KFR_AMD_API void* kfr_allocate_aligned(size_t size, size_t alignment);
KFR_AMD_API void kfr_deallocate(void* ptr);

/* complex DFT: include/kfr/capi.h:207-372 */
KFR_AMD_API KFR_DFT_PLAN_F32* kfr_dft_create_plan_f32(size_t size);
KFR_AMD_API KFR_DFT_PLAN_F64* kfr_dft_create_plan_f64(size_t size);
KFR_AMD_API KFR_DFT_PLAN_F32* kfr_dft_create_2d_plan_f32(size_t size1, size_t size2);
KFR_AMD_API KFR_DFT_PLAN_F64* kfr_dft_create_2d_plan_f64(size_t size1, size_t size2);
KFR_AMD_API KFR_DFT_PLAN_F32* kfr_dft_create_3d_plan_f32(size_t size1, size_t size2, size_t size3);
KFR_AMD_API KFR_DFT_PLAN_F64* kfr_dft_create_3d_plan_f64(size_t size1, size_t size2, size_t size3);
KFR_AMD_API KFR_DFT_PLAN_F32* kfr_dft_create_md_plan_f32(size_t dims, const unsigned* shape);
KFR_AMD_API KFR_DFT_PLAN_F64* kfr_dft_create_md_plan_f64(size_t dims, const unsigned* shape);
KFR_AMD_API void kfr_dft_dump_f32(KFR_DFT_PLAN_F32* plan);
KFR_AMD_API void kfr_dft_dump_f64(KFR_DFT_PLAN_F64* plan);
KFR_AMD_API size_t kfr_dft_get_size_f32(KFR_DFT_PLAN_F32* plan);
KFR_AMD_API size_t kfr_dft_get_size_f64(KFR_DFT_PLAN_F64* plan);
KFR_AMD_API size_t kfr_dft_get_temp_size_f32(KFR_DFT_PLAN_F32* plan);
KFR_AMD_API size_t kfr_dft_get_temp_size_f64(KFR_DFT_PLAN_F64* plan);
KFR_AMD_API void kfr_dft_execute_f32(KFR_DFT_PLAN_F32* plan, kfr_c32* out, const kfr_c32* in, uint8_t* temp);
KFR_AMD_API void kfr_dft_execute_f64(KFR_DFT_PLAN_F64* plan, kfr_c64* out, const kfr_c64* in, uint8_t* temp);
KFR_AMD_API void kfr_dft_execute_inverse_f32(KFR_DFT_PLAN_F32* plan, kfr_c32* out, const kfr_c32* in,
                                             uint8_t* temp);
KFR_AMD_API void kfr_dft_execute_inverse_f64(KFR_DFT_PLAN_F64* plan, kfr_c64* out, const kfr_c64* in,
                                             uint8_t* temp);
KFR_AMD_API void kfr_dft_delete_plan_f32(KFR_DFT_PLAN_F32* plan);
KFR_AMD_API void kfr_dft_delete_plan_f64(KFR_DFT_PLAN_F64* plan);

/* real DFT: include/kfr/capi.h:380-516 */
KFR_AMD_API KFR_DFT_REAL_PLAN_F32* kfr_dft_real_create_plan_f32(size_t size, KFR_DFT_PACK_FORMAT pack_format);
KFR_AMD_API KFR_DFT_REAL_PLAN_F64* kfr_dft_real_create_plan_f64(size_t size, KFR_DFT_PACK_FORMAT pack_format);
KFR_AMD_API KFR_DFT_REAL_PLAN_F32* kfr_dft_real_create_2d_plan_f32(size_t size1, size_t size2,
                                                                   kfr_bool real_out_is_enough);
KFR_AMD_API KFR_DFT_REAL_PLAN_F64* kfr_dft_real_create_2d_plan_f64(size_t size1, size_t size2,
                                                                   kfr_bool real_out_is_enough);
KFR_AMD_API KFR_DFT_REAL_PLAN_F32* kfr_dft_real_create_3d_plan_f32(size_t size1, size_t size2, size_t size3,
                                                                   kfr_bool real_out_is_enough);
KFR_AMD_API KFR_DFT_REAL_PLAN_F64* kfr_dft_real_create_3d_plan_f64(size_t size1, size_t size2, size_t size3,
                                                                   kfr_bool real_out_is_enough);
KFR_AMD_API KFR_DFT_REAL_PLAN_F32* kfr_dft_real_create_md_plan_f32(size_t dims, const unsigned* shape,
                                                                   kfr_bool real_out_is_enough);
KFR_AMD_API KFR_DFT_REAL_PLAN_F64* kfr_dft_real_create_md_plan_f64(size_t dims, const unsigned* shape,
                                                                   kfr_bool real_out_is_enough);
KFR_AMD_API void kfr_dft_real_dump_f32(KFR_DFT_REAL_PLAN_F32* plan);
KFR_AMD_API void kfr_dft_real_dump_f64(KFR_DFT_REAL_PLAN_F64* plan);
KFR_AMD_API size_t kfr_dft_real_get_size_f32(KFR_DFT_REAL_PLAN_F32* plan);
KFR_AMD_API size_t kfr_dft_real_get_size_f64(KFR_DFT_REAL_PLAN_F64* plan);
KFR_AMD_API size_t kfr_dft_real_get_temp_size_f32(KFR_DFT_REAL_PLAN_F32* plan);
KFR_AMD_API size_t kfr_dft_real_get_temp_size_f64(KFR_DFT_REAL_PLAN_F64* plan);
KFR_AMD_API void kfr_dft_real_execute_f32(KFR_DFT_REAL_PLAN_F32* plan, kfr_c32* out, const kfr_f32* in,
                                          uint8_t* temp);
KFR_AMD_API void kfr_dft_real_execute_f64(KFR_DFT_REAL_PLAN_F64* plan, kfr_c64* out, const kfr_f64* in,
                                          uint8_t* temp);
KFR_AMD_API void kfr_dft_real_execute_inverse_f32(KFR_DFT_REAL_PLAN_F32* plan, kfr_f32* out, const kfr_c32* in,
                                                  uint8_t* temp);
KFR_AMD_API void kfr_dft_real_execute_inverse_f64(KFR_DFT_REAL_PLAN_F64* plan, kfr_f64* out, const kfr_c64* in,
                                                  uint8_t* temp);
KFR_AMD_API void kfr_dft_real_delete_plan_f32(KFR_DFT_REAL_PLAN_F32* plan);
KFR_AMD_API void kfr_dft_real_delete_plan_f64(KFR_DFT_REAL_PLAN_F64* plan);

/* DCT-II / DCT-III: include/kfr/capi.h:523-636 */
KFR_AMD_API KFR_DCT_PLAN_F32* kfr_dct_create_plan_f32(size_t size);
KFR_AMD_API KFR_DCT_PLAN_F64* kfr_dct_create_plan_f64(size_t size);
KFR_AMD_API void kfr_dct_dump_f32(KFR_DCT_PLAN_F32* plan);
KFR_AMD_API void kfr_dct_dump_f64(KFR_DCT_PLAN_F64* plan);
KFR_AMD_API size_t kfr_dct_get_size_f32(KFR_DCT_PLAN_F32* plan);
KFR_AMD_API size_t kfr_dct_get_size_f64(KFR_DCT_PLAN_F64* plan);
KFR_AMD_API size_t kfr_dct_get_temp_size_f32(KFR_DCT_PLAN_F32* plan);
KFR_AMD_API size_t kfr_dct_get_temp_size_f64(KFR_DCT_PLAN_F64* plan);
KFR_AMD_API void kfr_dct_execute_f32(KFR_DCT_PLAN_F32* plan, kfr_f32* out, const kfr_f32* in, uint8_t* temp);
KFR_AMD_API void kfr_dct_execute_f64(KFR_DCT_PLAN_F64* plan, kfr_f64* out, const kfr_f64* in, uint8_t* temp);
KFR_AMD_API void kfr_dct_execute_inverse_f32(KFR_DCT_PLAN_F32* plan, kfr_f32* out, const kfr_f32* in,
                                             uint8_t* temp);
KFR_AMD_API void kfr_dct_execute_inverse_f64(KFR_DCT_PLAN_F64* plan, kfr_f64* out, const kfr_f64* in,
                                             uint8_t* temp);
KFR_AMD_API void kfr_dct_delete_plan_f32(KFR_DCT_PLAN_F32* plan);
KFR_AMD_API void kfr_dct_delete_plan_f64(KFR_DCT_PLAN_F64* plan);

/* filters: include/kfr/capi.h:644-736. Only the convolution filter (convolve_filter) is on the DFT
 * path; the FIR and IIR constructors return NULL with an error naming them as not provided. */
KFR_AMD_API KFR_FILTER_F32* kfr_filter_create_fir_plan_f32(const kfr_f32* taps, size_t size);
KFR_AMD_API KFR_FILTER_F64* kfr_filter_create_fir_plan_f64(const kfr_f64* taps, size_t size);
KFR_AMD_API KFR_FILTER_F32* kfr_filter_create_convolution_plan_f32(const kfr_f32* taps, size_t size,
                                                                   size_t block_size);
KFR_AMD_API KFR_FILTER_F64* kfr_filter_create_convolution_plan_f64(const kfr_f64* taps, size_t size,
                                                                   size_t block_size);
KFR_AMD_API KFR_FILTER_F32* kfr_filter_create_iir_plan_f32(const kfr_f32* sos, size_t sos_count);
KFR_AMD_API KFR_FILTER_F64* kfr_filter_create_iir_plan_f64(const kfr_f64* sos, size_t sos_count);
KFR_AMD_API void kfr_filter_process_f32(KFR_FILTER_F32* plan, kfr_f32* output, const kfr_f32* input, size_t size);
KFR_AMD_API void kfr_filter_process_f64(KFR_FILTER_F64* plan, kfr_f64* output, const kfr_f64* input, size_t size);
KFR_AMD_API void kfr_filter_reset_f32(KFR_FILTER_F32* plan);
KFR_AMD_API void kfr_filter_reset_f64(KFR_FILTER_F64* plan);
KFR_AMD_API void kfr_filter_delete_plan_f32(KFR_FILTER_F32* plan);
KFR_AMD_API void kfr_filter_delete_plan_f64(KFR_FILTER_F64* plan);

/* ============================== Part 2: extensions ============================== */

/* Batched transforms on device pointers, enqueued on `stream` (a hipStream_t; NULL = the default
 * stream) and not synchronised. Transforms are contiguous: item b starts at b*size (complex),
 * b*size reals / b*complex_size complex values (real), b*size reals (DCT). Returns 0 on success,
 * -1 on error (see kfr_last_error). */
KFR_AMD_API int kfr_amd_dft_execute_batch_f32(KFR_DFT_PLAN_F32* plan, kfr_c32* out, const kfr_c32* in,
                                              size_t batch, int inverse, void* stream);
KFR_AMD_API int kfr_amd_dft_execute_batch_f64(KFR_DFT_PLAN_F64* plan, kfr_c64* out, const kfr_c64* in,
                                              size_t batch, int inverse, void* stream);

/* kfr_amd_dft_execute_batch_f32 that allocates nothing: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The four-step and multi-pass intermediates live in one static 32 MiB device array
 * per GPU instead of the plan's scratch, in rounds that fit it. Refused with an error (return -1): one
 * item whose intermediates exceed 32 MiB (f32 four-step lengths above 2^22 points, the multi-pass
 * lengths from 2^24, mixed-radix products above 4,194,304 points) and Bluestein lengths whose transform
 * needs more than one workgroup. Calls that overlap on one GPU share the array and race; the plan's own
 * executions do not use it. */
KFR_AMD_API int kfr_amd_dft_execute_batch_noalloc_f32(KFR_DFT_PLAN_F32* plan, kfr_c32* out, const kfr_c32* in,
                                                      size_t batch, int inverse, void* stream);
KFR_AMD_API int kfr_amd_dft_real_execute_batch_f32(KFR_DFT_REAL_PLAN_F32* plan, void* out, const void* in,
                                                   size_t batch, int inverse, void* stream);
KFR_AMD_API int kfr_amd_dft_real_execute_batch_f64(KFR_DFT_REAL_PLAN_F64* plan, void* out, const void* in,
                                                   size_t batch, int inverse, void* stream);
KFR_AMD_API int kfr_amd_dct_execute_batch_f32(KFR_DCT_PLAN_F32* plan, kfr_f32* out, const kfr_f32* in,
                                              size_t batch, int inverse, void* stream);
KFR_AMD_API int kfr_amd_dct_execute_batch_f64(KFR_DCT_PLAN_F64* plan, kfr_f64* out, const kfr_f64* in,
                                              size_t batch, int inverse, void* stream);

/* Complex size of a real plan (size/2+1 for CCs, size/2 for Perm): dft_plan_real::complex_size,
 * include/kfr/dft/fft.hpp:471-476. */
KFR_AMD_API size_t kfr_amd_dft_real_get_complex_size_f32(KFR_DFT_REAL_PLAN_F32* plan);
KFR_AMD_API size_t kfr_amd_dft_real_get_complex_size_f64(KFR_DFT_REAL_PLAN_F64* plan);

/* convolve / correlate / autocorrelate (include/kfr/dft/convolution.hpp:53-76) of host or device
 * vectors. dtype: 0 f32, 1 f64, 2 c32, 3 c64; mode: 0 convolve, 1 correlate, 2 autocorrelate (b is
 * ignored). Writes the result to `out` when it is not NULL and returns its length, or -1 on error.
 * Complex inputs keep only the real part of the result, as the reference does. */
KFR_AMD_API long kfr_amd_convolve(int dtype, int mode, const void* a, size_t na, const void* b, size_t nb,
                                  void* out);

/* Batched real FFT convolution on device pointers with a hand-written adjoint.
 * y[b] = x[b] (*) h[b], |y| = nx + nh - 1; h_dist = 0 shares one filter across the batch.
 * backward: dx = g correlated with h (first nx), dh = g correlated with x (first nh, summed over the
 * batch when h_dist = 0); the signal spectrum is recomputed. dx or dh may be NULL. */
typedef struct KFR_AMD_CONV { int opaque; } KFR_AMD_CONV;
KFR_AMD_API KFR_AMD_CONV* kfr_amd_conv_create(int dtype, size_t nx, size_t nh);
KFR_AMD_API size_t kfr_amd_conv_fft_size(KFR_AMD_CONV* conv);
KFR_AMD_API int kfr_amd_conv_forward(KFR_AMD_CONV* conv, void* y, const void* x, const void* h, size_t h_dist,
                                     size_t batch, int correlate, void* stream);
KFR_AMD_API int kfr_amd_conv_backward(KFR_AMD_CONV* conv, void* dx, void* dh, const void* g, const void* x,
                                      const void* h, size_t h_dist, size_t batch, int correlate, void* stream);
KFR_AMD_API void kfr_amd_conv_delete(KFR_AMD_CONV* conv);

/* Uneven batches: `count` complex transforms of lengths sizes[i] and directions inverse[i] (NULL: all
 * forward) in one launch, as one persistent kernel fed from a device work queue and replayed as a
 * hipGraph. Transform i reads and writes sizes[0] + ... + sizes[i-1] complex values into the buffers.
 * Every length must fit one workgroup (f32 up to 8192 points, f64 up to 4096; lengths with a prime
 * factor of 101 or more run Bluestein inside the same kernel and may be up to half that). autotune != 0
 * times candidate schedules of each length on the device at creation: the order of its radix passes,
 * the LDS padding of its tile and the transforms per work item (cached per process). execute: device pointers are enqueued on `stream` (NULL = default stream) without
 * synchronising; host pointers are staged synchronously. Calls on one batch plan must not overlap.
 * dtype: 0 f32, 1 f64. No counterpart in the reference, which transforms one length per plan. */
typedef struct KFR_AMD_BATCH { int opaque; } KFR_AMD_BATCH;
KFR_AMD_API KFR_AMD_BATCH* kfr_amd_batch_create(int dtype, size_t count, const size_t* sizes, const int* inverse,
                                                int autotune);
KFR_AMD_API size_t kfr_amd_batch_total_size(KFR_AMD_BATCH* batch);
KFR_AMD_API int kfr_amd_batch_execute(KFR_AMD_BATCH* batch, void* out, const void* in, void* stream);
KFR_AMD_API int kfr_amd_batch_describe(KFR_AMD_BATCH* batch, char* buf, size_t buflen);
KFR_AMD_API void kfr_amd_batch_delete(KFR_AMD_BATCH* batch);

/* Complex transforms of power-of-two size stored in half precision: each value is a pair (re, im) of
 * bfloat16 (storage 1) or IEEE binary16 (storage 2), 4 bytes per complex value; arithmetic is FP32 and
 * the radix-16 passes run as 16-point DFT-matrix products on MFMA with the matrix and the data split
 * into bf16 hi + lo parts (FP32 accumulation). Transforms longer than one workgroup keep their
 * intermediate in binary16 with a power-of-two scale per group of columns. Unscaled, in place allowed,
 * contiguous batches; device pointers run on `stream` without synchronising, host pointers are staged
 * synchronously. No counterpart in the reference (whose transforms are float and double only). */
typedef struct KFR_AMD_DFT_HALF { int opaque; } KFR_AMD_DFT_HALF;
KFR_AMD_API KFR_AMD_DFT_HALF* kfr_amd_dft_half_create(int storage, size_t size);
KFR_AMD_API int kfr_amd_dft_half_execute_batch(KFR_AMD_DFT_HALF* plan, void* out, const void* in, size_t batch,
                                               int inverse, void* stream);
KFR_AMD_API int kfr_amd_dft_half_describe(KFR_AMD_DFT_HALF* plan, char* buf, size_t buflen);
KFR_AMD_API void kfr_amd_dft_half_delete(KFR_AMD_DFT_HALF* plan);

/* The reference plan descriptor for a size, computed on the host without a GPU: text in the layout of
 * dft_plan::dump() (one line per stage) preceded by a summary line. kind: 0 dft_plan, 1 dft_plan_real
 * (arg = pack format), 2 dct_plan, 3 dft_plan with progressive_optimized. Returns the text length
 * (the buffer receives at most buflen-1 bytes) or -1. */
KFR_AMD_API int kfr_amd_describe(int kind, int dtype, size_t size, int arg, char* buf, size_t buflen);

/* The permutation applied by the reference plan's reorder stage (out[i] = in[perm[i]]): radix-2 bit
 * reversal or mixed-radix digit reversal. Writes `size` indexes; returns 1, or 0 if the plan has no
 * reorder stage, or -1 on error. */
KFR_AMD_API int kfr_amd_dft_permutation(int dtype, size_t size, int64_t* perm);

/* Number of stages of the reference plan, i.e. dft_plan::progressive_total_steps()
 * (include/kfr/dft/fft.hpp:389-391). */
KFR_AMD_API size_t kfr_amd_dft_progressive_total_steps(int dtype, size_t size);

/* Plan cache: the process-wide cache of dft_cache (include/kfr/dft/cache.hpp:44-126: get(ctype, size),
 * getreal(ctype, size), clear()), keyed on element type, transform kind, size, pack format, batch
 * layout and schedule variant rather than on type and size only. Each call returns a new reference to
 * the cached plan, released with the matching kfr_dft_delete_plan_* / kfr_dft_real_delete_plan_*; the
 * plan stays valid until then, also across kfr_amd_cache_clear (which drops the cache's own
 * references). The convolutions (kfr_amd_convolve) take their plans from the same cache, as the
 * reference's convolve does (src/dft/convolution-impl.cpp:106-128). */
KFR_AMD_API KFR_DFT_PLAN_F32* kfr_amd_cache_dft_plan_f32(size_t size);
KFR_AMD_API KFR_DFT_PLAN_F64* kfr_amd_cache_dft_plan_f64(size_t size);
KFR_AMD_API KFR_DFT_REAL_PLAN_F32* kfr_amd_cache_dft_real_plan_f32(size_t size, KFR_DFT_PACK_FORMAT fmt);
KFR_AMD_API KFR_DFT_REAL_PLAN_F64* kfr_amd_cache_dft_real_plan_f64(size_t size, KFR_DFT_PACK_FORMAT fmt);
KFR_AMD_API void kfr_amd_cache_clear(void);
KFR_AMD_API size_t kfr_amd_cache_size(void);

/* convolve_filter<complex<T>> (include/kfr/dft/convolution.hpp:99-148 with a complex T; the reference's
 * C API declares the KFR_FILTER_C32/C64 handles, capi.h:163-164, but no creator): streaming FFT
 * convolution of complex signals with complex taps, the same contract as the real
 * kfr_filter_*_convolution functions. set_data replaces the taps and resets the state, as
 * convolve_filter::set_data does (src/dft/convolution-impl.cpp:58-84); the _f32/_f64 forms act on the
 * real filters of kfr_filter_create_convolution_plan_*. */
KFR_AMD_API KFR_FILTER_C32* kfr_amd_filter_create_convolution_plan_c32(const kfr_c32* taps, size_t size,
                                                                      size_t block_size);
KFR_AMD_API KFR_FILTER_C64* kfr_amd_filter_create_convolution_plan_c64(const kfr_c64* taps, size_t size,
                                                                      size_t block_size);
KFR_AMD_API void kfr_amd_filter_process_c32(KFR_FILTER_C32* plan, kfr_c32* output, const kfr_c32* input, size_t size);
KFR_AMD_API void kfr_amd_filter_process_c64(KFR_FILTER_C64* plan, kfr_c64* output, const kfr_c64* input, size_t size);
KFR_AMD_API void kfr_amd_filter_reset_c32(KFR_FILTER_C32* plan);
KFR_AMD_API void kfr_amd_filter_reset_c64(KFR_FILTER_C64* plan);
KFR_AMD_API void kfr_amd_filter_delete_plan_c32(KFR_FILTER_C32* plan);
KFR_AMD_API void kfr_amd_filter_delete_plan_c64(KFR_FILTER_C64* plan);
KFR_AMD_API void kfr_amd_filter_set_data_f32(KFR_FILTER_F32* plan, const kfr_f32* taps, size_t size);
KFR_AMD_API void kfr_amd_filter_set_data_f64(KFR_FILTER_F64* plan, const kfr_f64* taps, size_t size);
KFR_AMD_API void kfr_amd_filter_set_data_c32(KFR_FILTER_C32* plan, const kfr_c32* taps, size_t size);
KFR_AMD_API void kfr_amd_filter_set_data_c64(KFR_FILTER_C64* plan, const kfr_c64* taps, size_t size);

/* Progressive execution (dft_plan::progressive_start / progressive_step / progressive_total_steps,
 * include/kfr/dft/fft.hpp:393-439, src/dft/fft-impl.hpp:2025-2064; dft_plan_real's progressive_start,
 * fft.hpp:538-556). start records the buffers (host or device pointers) and returns a handle that keeps
 * the plan alive; each kfr_amd_progressive_step runs the next step synchronously and returns 1 while
 * steps remain, 0 after the last one (then `out` holds the result), -1 on error. The step count is the
 * reference plan's stage count for that direction (progressive_optimized as in dft_plan's constructor);
 * the transform's GPU launches (one Stockham pass each for lengths in one workgroup, the two kernels
 * of a four-step, the whole transform for Bluestein, plus the real repack or unpack) are spread over
 * the steps, the input is read at the first step and the output written at the last. */
typedef struct KFR_AMD_PROGRESSIVE { int opaque; } KFR_AMD_PROGRESSIVE;
KFR_AMD_API KFR_AMD_PROGRESSIVE* kfr_amd_dft_progressive_start_f32(KFR_DFT_PLAN_F32* plan, int progressive_optimized,
                                                                   int inverse, kfr_c32* out, const kfr_c32* in);
KFR_AMD_API KFR_AMD_PROGRESSIVE* kfr_amd_dft_progressive_start_f64(KFR_DFT_PLAN_F64* plan, int progressive_optimized,
                                                                   int inverse, kfr_c64* out, const kfr_c64* in);
KFR_AMD_API KFR_AMD_PROGRESSIVE* kfr_amd_dft_real_progressive_start_f32(KFR_DFT_REAL_PLAN_F32* plan, int inverse,
                                                                        void* out, const void* in);
KFR_AMD_API KFR_AMD_PROGRESSIVE* kfr_amd_dft_real_progressive_start_f64(KFR_DFT_REAL_PLAN_F64* plan, int inverse,
                                                                        void* out, const void* in);
KFR_AMD_API size_t kfr_amd_progressive_total_steps(KFR_AMD_PROGRESSIVE* prog);
KFR_AMD_API int kfr_amd_progressive_step(KFR_AMD_PROGRESSIVE* prog);
KFR_AMD_API void kfr_amd_progressive_delete(KFR_AMD_PROGRESSIVE* prog);
/* dft_plan<T>(size, normal, progressive_optimized).progressive_total_steps() and the same for
 * dft_plan_real<T>(size, fmt), computed on the host */
KFR_AMD_API size_t kfr_amd_dft_progressive_steps(int dtype, size_t size, int progressive_optimized);
KFR_AMD_API size_t kfr_amd_dft_real_progressive_steps(int dtype, size_t size, KFR_DFT_PACK_FORMAT fmt);

/* The GPU schedule chosen for a complex DFT size (algorithm, radices, launch geometry). */
KFR_AMD_API int kfr_amd_dft_schedule(int dtype, size_t size, char* buf, size_t buflen);

/* Test hook: the n-th following device allocation made by the library fails (0 disables). */
KFR_AMD_API void kfr_amd_inject_alloc_failure(int n);

/* Synchronise the library's stream. */
KFR_AMD_API int kfr_amd_synchronize(void);

/* Kernel timing. Between kfr_amd_timing_begin and kfr_amd_timing_end, every launch of a transform
 * kernel (power-of-two, mixed-radix, generic LDS, Bluestein and real-transform kernels) is bracketed by
 * two HIP events recorded on the stream it is launched on. kfr_amd_timing_end stops recording, waits
 * for those events and returns the number of launches recorded (-1 on error); kfr_amd_timing_get then
 * reads launch `index`: the kernel's template instance as profilers spell it (e.g.
 * "k_pow2<float, 12, false, 2, 0>"), its algorithmic bytes (one read of the values it transforms and
 * one write of its results, tables excluded) and its duration in milliseconds. Returns 0, or -1 for an
 * index out of range. No counterpart in the reference: a measurement aid. */
KFR_AMD_API int kfr_amd_timing_begin(void);
KFR_AMD_API long long kfr_amd_timing_end(void);
KFR_AMD_API int kfr_amd_timing_get(long long index, char* name, size_t namelen, long long* bytes, double* ms);

#ifdef __cplusplus
}
#endif
