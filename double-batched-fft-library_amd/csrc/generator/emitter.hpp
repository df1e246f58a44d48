// Internal entry points of the HIP source generator.
#ifndef BBFFT_AMD_EMITTER_HPP
#define BBFFT_AMD_EMITTER_HPP

#include "bbfft/detail/generator_impl.hpp"

#include <cstddef>
#include <iosfwd>
#include <string>

namespace bbfft {

/// Complex helper functions shared by all kernels of a source (include-guarded).
std::string hip_prelude();
/// One tile of cfg's kernel as `static __device__ void name(in, out, K, tile, lds)` for a
/// work-group of `group_size` work-items (>= cfg.work_group_size()).
void generate_fft1d_body(std::ostream &os, fft1d_configuration const &cfg,
                         std::string const &name, std::size_t group_size);
/// Tile body whose input (alias_in) or output (alias_out) pointer may point into the LDS
/// scratch passed to the body (real transforms included). lds_tables: the body reads its
/// twiddles from LDS copies that `name_tables(lid, n)` fills (emitted with the body).
void generate_fft1d_body_aliased(std::ostream &os, fft1d_configuration const &cfg,
                                 std::string const &name, std::size_t group_size, bool alias_in,
                                 bool alias_out, bool lds_tables = false);

} // namespace bbfft

#endif // BBFFT_AMD_EMITTER_HPP
