// Decomposition of a configuration into the 1-D double-batched passes that execute it.
#ifndef BBFFT_AMD_DECOMPOSE_HPP
#define BBFFT_AMD_DECOMPOSE_HPP

#include "bbfft/configuration.hpp"

#include <cstddef>
#include <vector>

namespace bbfft {

struct pass_layout {
    std::vector<configuration> passes; ///< 1-D configurations in execution order
    std::size_t tmp_bytes = 0;         ///< scratch buffer needed between passes (0: use out)
};

/**
 * @brief 1-D: one pass. 2-D/3-D: one double-batched pass per FFT mode, with the modes before
 * the transformed one folded into M and the modes after it folded into K, so no pass transposes.
 *
 * Reference: src/common/algorithm/nd_fft.hpp:25-112 (default layouts only, no callbacks; c2r
 * runs the passes in reverse order with swapped strides; a scratch buffer exists only when the
 * input tensor is larger than the output tensor).
 */
pass_layout decompose(configuration const &cfg);

} // namespace bbfft

#endif // BBFFT_AMD_DECOMPOSE_HPP
