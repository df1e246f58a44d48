// Lookup into the persisted autotuning table.
#ifndef BBFFT_AMD_TUNING_TABLE_HPP
#define BBFFT_AMD_TUNING_TABLE_HPP

#include "bbfft/configuration.hpp"
#include "bbfft/detail/generator_impl.hpp"

#include <cstddef>

namespace bbfft::gen {

/// Key bit of the entries tuned for variants of the persistent ragged kernel (type
/// "c2c-ragged" in the table): factorisation, work-items per transform, M tile, LDS padding
/// and access hints measured under that kernel's constraints.
constexpr unsigned tuning_ragged = 1u << 30;

/// Tuned parameters for (precision, type, N, M) and the input storage bits of `variant`
/// (half-precision input, MFMA stage), if the table has an entry. Only the access-hint flags
/// are stored; the caller adds the variant bits.
bool lookup_tuning(precision fp, transform_type type, std::size_t N, std::size_t M,
                   unsigned variant, fft1d_tuning &out);
std::size_t tuning_table_size();

} // namespace bbfft::gen

#endif // BBFFT_AMD_TUNING_TABLE_HPP
