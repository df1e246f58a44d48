// Scratch array for bbfft_plan_execute_static_scratch (defined in static_scratch.hip).
#ifndef BBFFT_AMD_STATIC_SCRATCH_HPP
#define BBFFT_AMD_STATIC_SCRATCH_HPP

#include "bbfft/bad_configuration.hpp"

#include <cstddef>

namespace bbfft::hip_detail {

/// Size of the per-GPU scratch array: 64 MiB.
constexpr std::size_t static_scratch_bytes = std::size_t(64) << 20;

/// Device address of `device`'s copy of the scratch array. Resolved once per device with
/// hipGetSymbolAddress (that device current during the call); allocates no device memory.
void *static_scratch(int device);

} // namespace bbfft::hip_detail

#endif // BBFFT_AMD_STATIC_SCRATCH_HPP
