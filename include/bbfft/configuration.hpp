// FFT configuration: shape, strides, precision, direction and transform type.
// Mirrors the reference interface include/bbfft/configuration.hpp:22-192.
#ifndef BBFFT_CONFIGURATION_HPP
#define BBFFT_CONFIGURATION_HPP

#include "bbfft/export.hpp"
#include "bbfft/user_module.hpp"

#include <array>
#include <cstddef>
#include <iosfwd>
#include <string>

namespace bbfft {

/// Floating-point precision; the value is the number of bytes of one real number.
enum class precision : int {
    f32 = 4, ///< single
    f64 = 8  ///< double
};

template <typename T> struct to_precision;
template <> struct to_precision<float> {
    static constexpr precision value = precision::f32;
};
template <> struct to_precision<double> {
    static constexpr precision value = precision::f64;
};
template <typename T> inline constexpr precision to_precision_v = to_precision<T>::value;

/// Sign of the exponent of the DFT kernel: forward is -1, backward is +1.
enum class direction : int {
    forward = -1, ///< exp(-2 pi i jk/N)
    backward = 1  ///< exp(+2 pi i jk/N)
};

/// Complex or real input/output.
enum class transform_type : int {
    c2c, ///< complex input, complex output
    r2c, ///< real input, complex output
    c2r  ///< complex input, real output
};
BBFFT_EXPORT char const *to_string(transform_type type);

constexpr unsigned max_fft_dim = 3;                 ///< largest FFT dimension
constexpr unsigned max_tensor_dim = max_fft_dim + 2; ///< FFT modes plus the M and K batch modes

/**
 * @brief Default strides of the input tensor M x N_1 x ... x N_d x K (column-major).
 *
 * s_0 = 1, s_1 = M, s_{i+1} = s_i N'_i where N'_1 is N_1 for c2c and out-of-place r2c,
 * 2(floor(N_1/2)+1) for in-place r2c, and floor(N_1/2)+1 for c2r.
 */
BBFFT_EXPORT auto default_istride(unsigned dim,
                                  std::array<std::size_t, max_tensor_dim> const &shape,
                                  transform_type type, bool inplace)
    -> std::array<std::size_t, max_tensor_dim>;

/// Default strides of the output tensor: the input rule with r2c and c2r swapped.
BBFFT_EXPORT auto default_ostride(unsigned dim,
                                  std::array<std::size_t, max_tensor_dim> const &shape,
                                  transform_type type, bool inplace)
    -> std::array<std::size_t, max_tensor_dim>;

/**
 * @brief Unified configuration for c2c/r2c/c2r transforms of dimension 1 to 3.
 *
 * Offsets are m s_0 + sum_i n_i s_i + k s_{d+1}, measured in reals for real tensors and in
 * complex numbers for complex tensors. s_0 must be 1.
 */
struct BBFFT_EXPORT configuration {
    unsigned dim;                                    ///< FFT dimension (1..max_fft_dim)
    std::array<std::size_t, max_tensor_dim> shape;   ///< {M, N_1, ..., N_d, K}
    precision fp;                                    ///< precision
    direction dir = direction::forward;              ///< direction
    transform_type type = transform_type::c2c;       ///< transform type
    std::array<std::size_t, max_tensor_dim> istride = default_istride(dim, shape, type, true);
    std::array<std::size_t, max_tensor_dim> ostride = default_ostride(dim, shape, type, true);
    user_module callbacks = {}; ///< optional load/store callbacks

    /// Set istride and ostride to the default layout.
    void set_strides_default(bool inplace);
    /// FFT descriptor string of this configuration.
    std::string to_string() const;
};

/// Writes the FFT descriptor of cfg.
BBFFT_EXPORT std::ostream &operator<<(std::ostream &os, configuration const &cfg);

} // namespace bbfft

#endif // BBFFT_CONFIGURATION_HPP
