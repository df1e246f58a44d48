// Tiling heuristics shared by the plain and the ragged kernel configuration.
#ifndef BBFFT_AMD_HEURISTICS_HPP
#define BBFFT_AMD_HEURISTICS_HPP

#include "bbfft/user_module.hpp"

#include <cstddef>
#include <cstdint>
#include <vector>

namespace bbfft::gen {

/// FNV-1a hash of a user module's source bytes and language (0 when it holds no callback).
std::uint64_t callback_hash(user_module const &um);

/// Stage radices for a complex FFT of length n: one stage up to `single_max`, otherwise the
/// fewest stages whose largest radix does not exceed `radix_max`, largest radix first.
std::vector<int> default_factorization(std::size_t n, std::size_t single_max,
                                       std::size_t radix_max);

} // namespace bbfft::gen

#endif // BBFFT_AMD_HEURISTICS_HPP
