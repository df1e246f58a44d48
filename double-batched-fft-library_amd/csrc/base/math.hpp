// Integer helpers used by the tiling and factorisation logic.
#ifndef BBFFT_AMD_MATH_HPP
#define BBFFT_AMD_MATH_HPP

#include <cstddef>
#include <vector>

namespace bbfft {

/// base^exp for small non-negative exponents (reference: src/base/math.cpp:8-17).
int ipow(int base, unsigned exp);
/// floor(radicand^(1/index)) (reference: src/base/math.cpp:19-40).
unsigned iroot(unsigned radicand, unsigned index);
/// Primality by trial division (reference: src/base/math.cpp:42-55).
bool is_prime(unsigned n);
/// Largest power of two <= x (1 for x < 2) (reference: src/base/math.cpp:57-63).
std::size_t max_power_of_2_less_equal(std::size_t x);
/// Smallest power of two >= x (reference: src/base/math.cpp:65-71).
std::size_t min_power_of_2_greater_equal(std::size_t x);
/// Prime factors of n in ascending order (reference: src/base/prime_factorization.cpp:13-26).
std::vector<int> trial_division(int n);
/**
 * @brief Split n into `index` factors whose product is n and which are closest, in the
 * least-squares sense, to n^(1/index); returned in ascending order.
 *
 * Reference: src/base/prime_factorization.cpp:28-77 (factor / update_factor).
 */
std::vector<unsigned> factor(unsigned n, unsigned index);

} // namespace bbfft

#endif // BBFFT_AMD_MATH_HPP
