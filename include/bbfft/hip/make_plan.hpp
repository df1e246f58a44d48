// make_plan for HIP streams.
// Mirrors the reference's include/bbfft/{cl,ze,sycl}/make_plan.hpp; like the Level Zero back end
// the plan uses caller-owned events (hipEvent_t).
#ifndef BBFFT_HIP_MAKE_PLAN_HPP
#define BBFFT_HIP_MAKE_PLAN_HPP

#include "bbfft/configuration.hpp"
#include "bbfft/export.hpp"
#include "bbfft/jit_cache.hpp"
#include "bbfft/plan.hpp"
#include "bbfft/ragged.hpp"

#include <hip/hip_runtime_api.h>

namespace bbfft {

using hip_plan = plan_unmanaged_event<hipEvent_t>;

/**
 * @brief Create a plan for cfg on the current HIP device; kernels run on `stream`.
 *
 * Kernels come from `cache` when present there, else they are generated and compiled with
 * hiprtc (and stored into `cache`). Throws bad_configuration for unsupported configurations and
 * bbfft::hip::error for runtime/compiler failures.
 */
BBFFT_EXPORT auto make_plan(configuration const &cfg, hipStream_t stream,
                            jit_cache *cache = nullptr) -> hip_plan;

/**
 * @brief Plan of a 1-D c2c fp32 configuration whose input is stored in half precision
 * (extension): `flags` is fft1d_input_f16 or fft1d_input_bf16, optionally with fft1d_mfma
 * (radix-16 first stage on the matrix cores). The output is fp32 complex.
 */
BBFFT_EXPORT auto make_half_input_plan(configuration const &cfg, unsigned flags,
                                       hipStream_t stream, jit_cache *cache = nullptr)
    -> hip_plan;

/**
 * @brief Plan for a ragged batch (extension): all jobs in one persistent kernel launch.
 *
 * execute(in, out) runs every job on its slice of in / out. With use_graph the launch is
 * captured into a HIP graph (once per distinct (in, out) pair) and replayed.
 */
BBFFT_EXPORT auto make_ragged_plan(ragged_configuration const &cfg, hipStream_t stream,
                                   jit_cache *cache = nullptr, bool use_graph = true)
    -> hip_plan;

} // namespace bbfft

#endif // BBFFT_HIP_MAKE_PLAN_HPP
