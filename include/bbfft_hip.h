/*
 * C ABI of the MI355X (gfx950) double-batched FFT library.
 *
 * Every entry point returns an int status (BBFFT_SUCCESS = 0) and never lets a C++ exception
 * cross the boundary; bbfft_last_error() returns the message of the last failure on the calling
 * thread. Streams and events are hipStream_t / hipEvent_t passed as opaque pointers, and tensors
 * are device pointers from hipMalloc (or any allocator returning device-accessible memory).
 *
 * Each declaration cites the reference interface it replaces (paths relative to the reference
 * repository intel/double-batched-fft-library).
 */
#ifndef BBFFT_HIP_H
#define BBFFT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BBFFT_EXPORT_C __attribute__((visibility("default")))

enum bbfft_status {
    BBFFT_SUCCESS = 0,
    BBFFT_ERROR_BAD_CONFIGURATION = 1, /* bbfft::bad_configuration */
    BBFFT_ERROR_RUNTIME = 2,           /* HIP runtime / hiprtc failure (bbfft::hip::error) */
    BBFFT_ERROR_INVALID_ARGUMENT = 3,  /* null handle, short buffer, index out of range */
    BBFFT_ERROR_PARSE = 4,             /* malformed descriptor (std::runtime_error in C++) */
    BBFFT_ERROR_INTERNAL = 5
};

typedef struct bbfft_plan_s *bbfft_plan_t;
typedef struct bbfft_cache_s *bbfft_cache_t;
typedef void *bbfft_stream_t; /* hipStream_t */
typedef void *bbfft_event_t;  /* hipEvent_t */

/*
 * Mirror of bbfft::configuration (include/bbfft/configuration.hpp:151-192) and
 * bbfft::user_module (include/bbfft/user_module.hpp:16-37).
 * precision: 4 (f32) or 8 (f64); direction: -1 forward, +1 backward;
 * type: 0 c2c, 1 r2c, 2 c2r; callback_language: 0 OpenCL-C style, 1 HIP.
 */
typedef struct bbfft_configuration {
    unsigned dim;
    size_t shape[5];
    int precision;
    int direction;
    int type;
    size_t istride[5];
    size_t ostride[5];
    const char *callback_source;
    size_t callback_length;
    const char *load_function;
    const char *store_function;
    int callback_language;
} bbfft_configuration;

/* Tiling of one generated kernel (include/bbfft/detail/generator_impl.hpp: fft1d_configuration;
 * replaces the reference's small_batch_configuration / factor2_slm_configuration,
 * include/bbfft/detail/generator_impl.hpp:26-107 of the reference). */
typedef struct bbfft_kernel_config {
    size_t M, Mb, N, K, Nt, Kb, lds_pad, lds_bytes, work_group_size;
    int factorization[8];
    int num_factors;
    int inplace_unsupported;
} bbfft_kernel_config;

/* Overrides of the tiling heuristics of a 1-D plan (0 / num_factors == 0: automatic).
 * flags: bit 0 non-temporal global loads, bit 1 non-temporal global stores, BBFFT_PREFETCH
 * (c2c / r2c without callbacks: each work-group loads its next tile's inputs while it transforms
 * the current one; for persistent grids, groups_per_cu > 0); used only when has_flags != 0.
 * (No reference counterpart: the reference's tiling is fixed by device_info.) */
enum { BBFFT_PREFETCH = 64 };
typedef struct bbfft_tuning {
    int factorization[8];
    int num_factors;
    size_t Nt, Mb, Kb, lds_pad;
    int no_pad;
    unsigned flags;
    int has_flags;
    unsigned waves_per_eu; /* occupancy target of the compiler, 0..8 (0 = its default) */
    unsigned groups_per_cu; /* persistent grid: work-groups per CU (0 = one per tile) */
} bbfft_tuning;

/* Message of the last failed call on this thread ("" if none). */
BBFFT_EXPORT_C const char *bbfft_last_error(void);
/* Library version string. */
BBFFT_EXPORT_C const char *bbfft_version(void);

/* parse_fft_descriptor (src/base/parser.cpp:59-211, include/bbfft/parser.hpp:24). */
BBFFT_EXPORT_C int bbfft_parse_fft_descriptor(const char *desc, bbfft_configuration *cfg);
/* configuration::to_string / operator<< (src/base/configuration.cpp:63-160). */
BBFFT_EXPORT_C int bbfft_configuration_to_string(const bbfft_configuration *cfg, char *buf,
                                                 size_t len);
/* default_istride / default_ostride (src/base/configuration.cpp:24-56). */
BBFFT_EXPORT_C int bbfft_default_istride(unsigned dim, const size_t shape[5], int type,
                                         int inplace, size_t stride[5]);
BBFFT_EXPORT_C int bbfft_default_ostride(unsigned dim, const size_t shape[5], int type,
                                         int inplace, size_t stride[5]);
/* parse_device_info + device_info::to_string (src/base/parser.cpp:213-270,
 * src/base/device_info.cpp:59-95): parse `desc` and print it canonically into buf. */
BBFFT_EXPORT_C int bbfft_device_info_normalize(const char *desc, char *buf, size_t len);
/* Integer helpers of the factorisation logic (src/base/math.cpp:8-71,
 * src/base/prime_factorization.cpp:13-77): factor(n, index) writes `index` factors. */
BBFFT_EXPORT_C int bbfft_factor(unsigned n, unsigned index, unsigned *factors);
BBFFT_EXPORT_C unsigned bbfft_iroot(unsigned radicand, unsigned index);
BBFFT_EXPORT_C int bbfft_is_prime(unsigned n);

/* exp(2 pi i k / n) as the generator emits it: correctly rounded to double (precision 8) or
 * float (precision 4) (plays the role of power_of_w / w_lut_N, src/base/root_of_unity.cpp:144-189
 * and src/base/root_of_unity.hpp:16-1162). */
BBFFT_EXPORT_C int bbfft_root_of_unity(long k, long n, int precision, double *re, double *im);

/* Kernel configuration and HIP source of the 1-D pass `pass` of cfg for a device-info string
 * (configure_* + generate_*, include/bbfft/detail/generator_impl.hpp:51-107 of the reference).
 * No GPU is needed. bbfft_generate_source writes at most len bytes and sets *needed. */
BBFFT_EXPORT_C int bbfft_num_passes(const bbfft_configuration *cfg, unsigned *num);
BBFFT_EXPORT_C int bbfft_kernel_configuration(const bbfft_configuration *cfg, unsigned pass,
                                              const char *device_info,
                                              bbfft_kernel_config *out, char *name,
                                              size_t name_len);
BBFFT_EXPORT_C int bbfft_generate_source(const bbfft_configuration *cfg, unsigned pass,
                                         const char *device_info, char *buf, size_t len,
                                         size_t *needed);
/* Same as bbfft_kernel_configuration / bbfft_generate_source with explicit tuning (1-D). */
BBFFT_EXPORT_C int bbfft_kernel_configuration_tuned(const bbfft_configuration *cfg,
                                                    const bbfft_tuning *tuning,
                                                    const char *device_info,
                                                    bbfft_kernel_config *out, char *name,
                                                    size_t name_len);
BBFFT_EXPORT_C int bbfft_generate_source_tuned(const bbfft_configuration *cfg,
                                               const bbfft_tuning *tuning,
                                               const char *device_info, char *buf, size_t len,
                                               size_t *needed);

/* Compile the kernel of pass `pass` (callbacks included) with hiprtc for `arch`; no GPU needed.
 * The compiler log of a failure is returned by bbfft_last_error(). */
BBFFT_EXPORT_C int bbfft_compile_kernel(const bbfft_configuration *cfg, unsigned pass,
                                        const char *device_info, const char *arch,
                                        size_t *code_size);

/* Module caches: jit_cache_all (include/bbfft/jit_cache_all.hpp:20-37) layered over an
 * aot_cache (include/bbfft/aot_cache.hpp:21-48). bbfft_cache_load_aot registers an archive
 * written by bbfft_aot_generate (tools/aot/main.cpp:22-60 + aot_cache.register_module). */
BBFFT_EXPORT_C int bbfft_cache_create(bbfft_cache_t *cache);
BBFFT_EXPORT_C int bbfft_cache_load_aot(bbfft_cache_t cache, const char *path, int device,
                                        size_t *num_kernels);
/* Register an archive held in memory (linked into the executable: `python -m bbfft_amd.aot
 * --embed` writes it as a C++ array; the reference links archives with `ld -r -b binary`,
 * cmake/AddAotKernelsToTarget.cmake:4-33). */
BBFFT_EXPORT_C int bbfft_cache_load_aot_memory(bbfft_cache_t cache, const void *data, size_t size,
                                               int device, size_t *num_kernels);
/* Host-only listing of an archive in memory: one "arch kernel_name" line per kernel. */
BBFFT_EXPORT_C int bbfft_aot_archive_kernels(const void *data, size_t size, char *buf, size_t len,
                                             size_t *needed);
BBFFT_EXPORT_C int bbfft_cache_size(bbfft_cache_t cache, size_t *num_modules);
BBFFT_EXPORT_C void bbfft_cache_destroy(bbfft_cache_t cache);

/* make_plan(cfg, queue, cache) (include/bbfft/ze/make_plan.hpp:29-31,
 * include/bbfft/sycl/make_plan.hpp:27-41): plan on the current HIP device for `stream`. */
BBFFT_EXPORT_C int bbfft_plan_create(const bbfft_configuration *cfg, bbfft_stream_t stream,
                                     bbfft_cache_t cache, bbfft_plan_t *plan);
BBFFT_EXPORT_C int bbfft_plan_create_from_descriptor(const char *desc, bbfft_stream_t stream,
                                                     bbfft_cache_t cache, bbfft_plan_t *plan);
/* Adjoint (transpose w.r.t. the real inner product) of cfg's transform, for differentiation.
 * No reference counterpart (an extension): c2c -> opposite-sign unnormalised transform with
 * in/out layouts swapped; r2c -> Hermitian-fold weighting + c2r; c2r -> r2c + doubling of the
 * non-self-conjugate bins. bbfft_adjoint_configuration returns the adjoint's configuration
 * (its real pass additionally runs the weighting variant). Callbacks of cfg apply to the
 * adjoint plan's own input (load) and output (store) where they run in a c2c pass (1-D c2c, the
 * first or last pass of 2-D/3-D plans); otherwise plan creation fails with
 * BBFFT_ERROR_BAD_CONFIGURATION. */
BBFFT_EXPORT_C int bbfft_adjoint_configuration(const bbfft_configuration *cfg,
                                               bbfft_configuration *adjoint);
BBFFT_EXPORT_C int bbfft_plan_create_adjoint(const bbfft_configuration *cfg,
                                             bbfft_stream_t stream, bbfft_cache_t cache,
                                             bbfft_plan_t *plan);
/* make_plan of a 1-D configuration with explicit tuning (autotuner candidates). */
BBFFT_EXPORT_C int bbfft_plan_create_tuned(const bbfft_configuration *cfg,
                                           const bbfft_tuning *tuning, bbfft_stream_t stream,
                                           bbfft_cache_t cache, bbfft_plan_t *plan);
/* plan_unmanaged_event::execute (include/bbfft/plan.hpp:165-183): waits for `wait`, launches,
 * records `signal` (both optional); in == out is the in-place transform. Asynchronous. */
BBFFT_EXPORT_C int bbfft_plan_execute(bbfft_plan_t plan, const void *in, void *out,
                                      bbfft_event_t signal, uint32_t num_wait,
                                      const bbfft_event_t *wait);
/* Allocation-free variant of bbfft_plan_execute. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). It calls no hipMalloc/hipFree and no allocating
 * library routine. Plans whose passes need a scratch buffer between them (2-D/3-D plans whose
 * input is larger than their output, i.e. out-of-place c2r) use a file-scope device array of
 * 64 MiB per GPU in place of the plan's own buffer; a plan needing more is rejected with
 * BBFFT_ERROR_BAD_CONFIGURATION before anything is launched. Ragged plans launch directly
 * (no graph capture). The array needs no clearing (every pass writes what the next reads),
 * keeps the last call's contents, and is shared by every call on that GPU: calls that overlap
 * on one GPU (e.g. on two streams) race on it. */
BBFFT_EXPORT_C int bbfft_plan_execute_static_scratch(bbfft_plan_t plan, const void *in,
                                                     void *out, bbfft_event_t signal,
                                                     uint32_t num_wait,
                                                     const bbfft_event_t *wait);
/* Introspection of the kernels a plan launches. */
BBFFT_EXPORT_C int bbfft_plan_num_launches(bbfft_plan_t plan, unsigned *num);
BBFFT_EXPORT_C int bbfft_plan_launch_info(bbfft_plan_t plan, unsigned index, char *name,
                                          size_t name_len, unsigned *grid, unsigned *block,
                                          size_t *lds_bytes);
BBFFT_EXPORT_C void bbfft_plan_destroy(bbfft_plan_t plan);

/* Device info of a HIP device as a device-info string (src/cl/device.cpp:12-40 pattern). */
BBFFT_EXPORT_C int bbfft_device_info_query(int device, char *buf, size_t len);

/*
 * Fused passes of 2-D/3-D plans (extension of the reference's nd_fft,
 * src/common/algorithm/nd_fft.hpp:52-99, which runs one memory round trip per mode): where a
 * plane of the intermediate tensor fits in LDS (up to 96 KB with at least 256 planes, or any
 * size when the problem is launch-bound), the passes over modes 1 and 2 run as one kernel.
 * Where a second buffer for one plane's input also fits, the kernel streams the next plane's
 * input into it during the current plane (lds_bytes below includes that buffer). These entry
 * points report the decision and the kernel for a device_info (host only); plans apply it
 * automatically (BBFFT_FUSE=off disables fusion, BBFFT_FUSE=on forces it where it fits,
 * BBFFT_FUSE_STAGE=off disables the input staging).
 */
typedef struct bbfft_fused_info {
    unsigned first_pass;
    size_t group_size, lds_bytes, planes, ka, kb, plane_elems, in_plane_stride,
        out_plane_stride;
} bbfft_fused_info;

BBFFT_EXPORT_C int bbfft_fused_configuration(const bbfft_configuration *cfg,
                                             const char *device_info, int *fused,
                                             bbfft_fused_info *info, char *name,
                                             size_t name_len);
BBFFT_EXPORT_C int bbfft_generate_fused_source(const bbfft_configuration *cfg,
                                               const char *device_info, char *buf, size_t len,
                                               size_t *needed);

/*
 * Half-precision input (extension; no reference counterpart). For a 1-D c2c fp32
 * configuration, the input holds (re, im) pairs of IEEE binary16 (BBFFT_INPUT_F16) or bfloat16
 * (BBFFT_INPUT_BF16) with the configuration's strides (counted in complex elements, 4 bytes
 * each); the output is fp32 complex. With BBFFT_MFMA the first stage (radix 16; N must be a
 * multiple of 16, N >= 32) runs as a DFT-matrix GEMM on the matrix cores with the matrix split
 * into 2 (f16) or 3 (bf16) half-precision terms and fp32 accumulation.
 */
enum { BBFFT_INPUT_F16 = 8, BBFFT_INPUT_BF16 = 16, BBFFT_MFMA = 32 };

/* Tiling a half-input plan uses (the tuning table included): `flags` as for
 * bbfft_plan_create_half_input. */
BBFFT_EXPORT_C int bbfft_half_input_kernel_configuration(const bbfft_configuration *cfg,
                                                         unsigned flags,
                                                         const char *device_info,
                                                         bbfft_kernel_config *out, char *name,
                                                         size_t name_len);
BBFFT_EXPORT_C int bbfft_plan_create_half_input(const bbfft_configuration *cfg, unsigned flags,
                                                bbfft_stream_t stream, bbfft_cache_t cache,
                                                bbfft_plan_t *plan);

/*
 * Ragged batches (extension; no reference counterpart: the reference runs one plan per shape,
 * src/common/algorithm/factor2_slm_fft.hpp:92-203). A job is a c2c M x N x K tensor with packed
 * column-major strides {1, M, M*N}; in_offset / out_offset count complex elements from the
 * pointers given to bbfft_plan_execute. precision: 4 or 8; direction: -1 forward, +1 backward.
 */
typedef struct bbfft_ragged_job {
    size_t M, N, K;
    int direction;
    size_t in_offset, out_offset;
} bbfft_ragged_job;

/* Persistent kernel of a ragged batch (include/bbfft/detail/generator_impl.hpp:
 * ragged_kernel_configuration). */
typedef struct bbfft_ragged_info {
    size_t group_size, lds_bytes, num_variants;
    unsigned long long total_tiles;
} bbfft_ragged_info;

enum { BBFFT_RAGGED_GRAPH = 1 }; /* replay the launch from a captured HIP graph */

BBFFT_EXPORT_C int bbfft_ragged_kernel_configuration(int precision, const bbfft_ragged_job *jobs,
                                                     size_t num_jobs, const char *device_info,
                                                     bbfft_ragged_info *info, char *name,
                                                     size_t name_len);
/* Tiling of variant `variant` (0 <= variant < num_variants); K is left 0. */
BBFFT_EXPORT_C int bbfft_ragged_variant_configuration(int precision,
                                                      const bbfft_ragged_job *jobs,
                                                      size_t num_jobs, const char *device_info,
                                                      size_t variant, bbfft_kernel_config *out);
/* Variant of each job (job_variant[num_jobs]). */
BBFFT_EXPORT_C int bbfft_ragged_job_variants(int precision, const bbfft_ragged_job *jobs,
                                             size_t num_jobs, const char *device_info,
                                             unsigned *job_variant);
/* Device job table for a launch of `grid` work-groups (5 * num_jobs + 2 * grid + 2 64-bit
 * words: tile starts, job records, per-work-group tile ranges), as the kernel reads it. */
BBFFT_EXPORT_C int bbfft_ragged_job_table(int precision, const bbfft_ragged_job *jobs,
                                          size_t num_jobs, const char *device_info, size_t grid,
                                          unsigned long long *table, size_t len, size_t *needed);
BBFFT_EXPORT_C int bbfft_generate_ragged_source(int precision, const bbfft_ragged_job *jobs,
                                                size_t num_jobs, const char *device_info,
                                                char *buf, size_t len, size_t *needed);
/* flags: BBFFT_RAGGED_GRAPH or 0. The plan is executed and destroyed like any other plan. */
BBFFT_EXPORT_C int bbfft_ragged_plan_create(int precision, const bbfft_ragged_job *jobs,
                                            size_t num_jobs, unsigned flags,
                                            bbfft_stream_t stream, bbfft_cache_t cache,
                                            bbfft_plan_t *plan);
/* Ahead-of-time compilation of the ragged kernel for these jobs' variants, plus the plain
 * kernel of every variant with a K tile of 1 (what bbfft_plan_create_tuned builds for the
 * variant's factorisation, Nt, Mb and lds_pad with Kb = 1 and flags 0: per-job launches with
 * the same arithmetic). The archive loads with bbfft_cache_load_aot. */
BBFFT_EXPORT_C int bbfft_aot_generate_ragged(int precision, const bbfft_ragged_job *jobs,
                                             size_t num_jobs, const char *device_info,
                                             const char *arch, const char *out_path);

/* bbfft-aot-generate (tools/aot/main.cpp:22-60): generate every kernel the descriptors need for
 * device_info, compile them with hiprtc for `arch` (e.g. "gfx950") and write an archive to
 * out_path. Kernel i of the de-duplicated list is compiled only when i % num_shards == shard, so
 * independent processes can build disjoint archives in parallel (hiprtc compiles one program
 * at a time per process). No GPU is needed. */
BBFFT_EXPORT_C int bbfft_aot_generate(const char *const *descriptors, size_t num_descriptors,
                                      const char *device_info, const char *arch,
                                      const char *out_path, unsigned shard, unsigned num_shards,
                                      size_t *num_kernels);

#ifdef __cplusplus
}
#endif

#endif /* BBFFT_HIP_H */
