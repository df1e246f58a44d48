// Decomposition of a configuration into the 1-D double-batched passes that execute it.
#ifndef BBFFT_AMD_DECOMPOSE_HPP
#define BBFFT_AMD_DECOMPOSE_HPP

#include "bbfft/configuration.hpp"
#include "bbfft/detail/generator_impl.hpp"
#include "bbfft/device_info.hpp"

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

/**
 * @brief Configuration of the adjoint transform (transpose w.r.t. the real inner product).
 *
 * c2c: the opposite direction with input and output strides swapped (the unnormalised DFT
 * matrix is symmetric, so its adjoint is its conjugate). r2c (c2r): a c2r (r2c) transform from
 * the output layout to the input layout; its real pass must run with the fft1d_adjoint variant
 * (Hermitian-fold weighting), which `real_pass_variant` reports.
 */
configuration adjoint_configuration(configuration const &cfg, unsigned *real_pass_variant);

/// Two consecutive passes of a layout that run as one fused kernel.
struct fused_pair {
    std::size_t first; ///< index of the first fused pass
    fused2_configuration fc;
};

/**
 * @brief The pair of passes of a 2-D/3-D layout to fuse on this device, if any: the first two
 * passes (modes 1 and 2) for c2c and r2c, the last two for c2r (the passes run in reverse
 * mode order), when the plane is small enough for several work-groups per CU or the whole
 * problem is launch-bound. `real_pass_variant` applies to the real pass. BBFFT_FUSE=off
 * disables fusion, BBFFT_FUSE=on applies it whenever the plane fits.
 */
bool find_fused_pair(pass_layout const &layout, device_info const &info,
                     unsigned real_pass_variant, fused_pair &out);

} // namespace bbfft

#endif // BBFFT_AMD_DECOMPOSE_HPP
