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
/// Options of a tile body in a fused kernel.
struct body_options {
    /// The body reads its twiddles from LDS copies that `name_tables(lid, n)` fills (emitted
    /// with the body).
    bool lds_tables = false;
    /// Barriers order LDS accesses only (BBFFT_LDS_SYNC(), which the kernel's source defines).
    bool lds_barriers = false;
    /// Extra parameter appended to the body's signature (e.g. "const float* next").
    std::string extra_param;
    /// Statement emitted once stage 0 has read all of the tile's input (after its barrier).
    std::string after_input;
    /// Expression XOR-ed into the element index of every input read (c2c, or r2c with
    /// even N read as complex pairs): the input sits in a row-wise permuted layout.
    std::string input_index_xor;
    /// The body is always called with tile 0 and K equal to its column tile (one tile covers
    /// all columns, M = Mb): every work-item of the group's band has a column, so `active` is
    /// the band alone and work-items outside it neither load nor compute.
    bool single_tile = false;
    /// With lds_tables: rows of R twiddles (one row per j mod p) sit R + 1 entries apart in
    /// LDS, so work-items reading different rows at once do not share banks.
    bool lds_table_pad = false;
};
/// Tile body whose input (alias_in) or output (alias_out) pointer may point into the LDS
/// scratch passed to the body (real transforms included).
void generate_fft1d_body_aliased(std::ostream &os, fft1d_configuration const &cfg,
                                 std::string const &name, std::size_t group_size, bool alias_in,
                                 bool alias_out, body_options const &opt = {});

} // namespace bbfft

#endif // BBFFT_AMD_EMITTER_HPP
