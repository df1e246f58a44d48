// Tiling heuristics shared by the plain and the ragged kernel configuration.
#ifndef BBFFT_AMD_HEURISTICS_HPP
#define BBFFT_AMD_HEURISTICS_HPP

#include <cstddef>
#include <vector>

namespace bbfft::gen {

/// Stage radices for a complex FFT of length n: one stage up to `single_max`, otherwise the
/// fewest stages whose largest radix does not exceed `radix_max`, largest radix first.
std::vector<int> default_factorization(std::size_t n, std::size_t single_max,
                                       std::size_t radix_max);

} // namespace bbfft::gen

#endif // BBFFT_AMD_HEURISTICS_HPP
