// Correctly rounded roots of unity for twiddle constants and tables.
#ifndef BBFFT_AMD_ROOTS_HPP
#define BBFFT_AMD_ROOTS_HPP

#include <complex>

namespace bbfft::gen {

/**
 * @brief exp(2 pi i k / n) rounded to double (to nearest).
 *
 * The angle is reduced exactly with integer arithmetic to [0, pi/4] and evaluated in 113-bit
 * quad precision, so values on the axes and diagonals are exact and conjugate/rotated pairs are
 * exactly symmetric. This plays the role of the reference's exactly rounded look-up tables
 * (src/base/root_of_unity.hpp:16-18, root_of_unity.cpp:144-189) for every n.
 */
std::complex<double> root_of_unity(long k, long n);
/// Same value rounded directly from quad precision to float (no double rounding).
std::complex<float> root_of_unity_f(long k, long n);

} // namespace bbfft::gen

#endif // BBFFT_AMD_ROOTS_HPP
