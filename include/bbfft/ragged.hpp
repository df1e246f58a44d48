// Ragged batches: many independent 1-D complex transforms of different length N, left batch M,
// right batch K and direction, computed by one persistent kernel launch.
//
// No counterpart in the reference, which runs one plan (one kernel) per (M, N, K) shape
// (src/common/algorithm/factor2_slm_fft.hpp:92-203); a ragged plan replaces a sequence of such
// plans whose shapes differ.
#ifndef BBFFT_RAGGED_HPP
#define BBFFT_RAGGED_HPP

#include "bbfft/configuration.hpp"
#include "bbfft/export.hpp"

#include <cstddef>
#include <vector>

namespace bbfft {

/**
 * @brief One job of a ragged batch: a c2c M x N x K tensor with the default (packed,
 * column-major) strides {1, M, M * N}.
 *
 * The job's input starts `in_offset` complex elements after the input pointer passed to
 * execute and its output `out_offset` complex elements after the output pointer. Jobs must not
 * overlap each other's output; a job may be in-place (same pointer and offset).
 */
struct ragged_job {
    std::size_t M = 1;
    std::size_t N = 1;
    std::size_t K = 1;
    direction dir = direction::forward;
    std::size_t in_offset = 0;
    std::size_t out_offset = 0;
};

/// A ragged batch: all jobs share the precision.
struct ragged_configuration {
    precision fp = precision::f32;
    std::vector<ragged_job> jobs;
};

} // namespace bbfft

#endif // BBFFT_RAGGED_HPP
