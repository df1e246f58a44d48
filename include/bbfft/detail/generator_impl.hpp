// Kernel-level configuration of the gfx950 generator and its entry points.
// Plays the role of the reference's include/bbfft/detail/generator_impl.hpp:26-107, where two
// kernel families (small_batch_configuration, factor2_slm_configuration) exist; on gfx950 one
// family covers both: a single radix stage owned by one work-item is the small-batch case and
// several stages exchanged through LDS are the multi-factor case.
#ifndef BBFFT_DETAIL_GENERATOR_IMPL_HPP
#define BBFFT_DETAIL_GENERATOR_IMPL_HPP

#include "bbfft/configuration.hpp"
#include "bbfft/device_info.hpp"
#include "bbfft/export.hpp"
#include "bbfft/ragged.hpp"

#include <array>
#include <cstddef>
#include <cstdint>
#include <iosfwd>
#include <string>
#include <string_view>
#include <vector>

namespace bbfft {

/**
 * @brief Tiling and algorithm of one generated 1-D double-batched FFT kernel.
 *
 * A work-group of Mb * Nt * Kb work-items owns Mb consecutive M-columns times Kb consecutive
 * K-columns (K-column pairs for real transforms of odd length). Nt work-items cooperate on one
 * transform: the complex FFT of length N_fft runs as len(factorization) Stockham stages, stage s
 * with radix factorization[s] in registers; stages exchange data through padded LDS.
 */
struct BBFFT_EXPORT fft1d_configuration {
    int direction;                       ///< -1 forward, +1 backward
    std::size_t M;                       ///< M (left batch)
    std::size_t Mb;                      ///< M tile
    std::size_t N;                       ///< transform length (real length for r2c/c2r)
    std::vector<int> factorization;      ///< stage radices of the complex FFT, first stage first
    std::size_t Nt;                      ///< work-items per transform
    std::size_t Kb;                      ///< K tile
    precision fp;                        ///< precision
    transform_type type;                 ///< c2c / r2c / c2r
    std::array<std::size_t, 3u> istride; ///< {1, s_N, s_K} of the input
    std::array<std::size_t, 3u> ostride; ///< {1, s_N, s_K} of the output
    std::size_t lds_pad;                 ///< one padding slot per lds_pad slots (0 = none)
    unsigned flags;                      ///< fft1d_flag bits (global access hints)
    unsigned waves_per_eu;               ///< occupancy target of the compiler (0 = default)
    unsigned groups_per_cu;              ///< persistent grid: work-groups per CU (0 = a tile each)
    bool inplace_unsupported;            ///< in == out would race across work-groups
    char const *load_function;           ///< user load callback or nullptr
    char const *store_function;          ///< user store callback or nullptr
    /// FNV-1a hash of the callback source and its language (0 without callbacks): two plans
    /// whose callbacks share names but not code get different identifiers and cache entries.
    std::uint64_t callback_hash = 0;

    /// Kernel symbol name; also the jit_cache key.
    std::string identifier() const;
    /// Length of the complex FFT (N/2 for real even N, else N).
    std::size_t N_fft() const;
    /// True for real transforms of odd length (two K-columns share one complex FFT).
    bool pairs_k() const;
    /// Number of K-columns (or K-column pairs) for a given K.
    std::size_t K_columns(std::size_t K) const;
    std::size_t work_group_size() const { return Mb * Nt * Kb; }
    /// Bytes of LDS used by one work-group (stage exchange plus direct-stage root tables).
    std::size_t lds_bytes() const;
    /// Radix of the direct DFT stage (a radix above fft1d_max_register_radix), 0 if none.
    int direct_radix() const;
};

/// Bits of fft1d_configuration::flags.
enum fft1d_flag : unsigned {
    fft1d_nontemporal_load = 1u,  ///< stream the input past the caches (read once)
    fft1d_nontemporal_store = 2u, ///< stream the output past the caches (written once)
    /// Real transforms only: the kernel computes the adjoint (transpose w.r.t. the real inner
    /// product) of the opposite real transform. c2r: the Hermitian-fold weights (1/2 on the
    /// non-self-conjugate bins, Nyquist imaginary part dropped) are applied in the pre-pass,
    /// making it the adjoint of r2c. r2c: non-self-conjugate bins are doubled and, for even N,
    /// imag(X[N/2]) = -real(X[0]) in the post-pass, making it the adjoint of c2r (including the
    /// c2r Nyquist rule).
    fft1d_adjoint = 4u,
    /// c2c fp32 only: the input is stored as IEEE half-precision complex (re, im) pairs (4 bytes
    /// per element, real part in the low half); the output stays fp32 complex.
    fft1d_input_f16 = 8u,
    /// As fft1d_input_f16 with bfloat16 (re, im) pairs.
    fft1d_input_bf16 = 16u,
    /// With fft1d_input_f16 / fft1d_input_bf16: the first stage (radix 16) runs as a DFT-matrix
    /// GEMM on the matrix cores (v_mfma_f32_32x32x16_{f16,bf16}) with fp32 accumulation; the
    /// DFT matrix is split into 2 (f16) or 3 (bf16) half-precision terms whose products with the
    /// exactly representable input are accumulated, so the stage is accurate to fp32 level.
    fft1d_mfma = 32u,
    /// Software-pipelined tile loop: while a work-group transforms one tile it already loads the
    /// first-stage inputs of its next tile (tile + number of work-groups) into registers, so
    /// the memory latency of a tile overlaps the LDS stages and stores of the previous one.
    /// Pays off with a persistent grid (groups_per_cu > 0). c2c and r2c without callbacks;
    /// not with fft1d_mfma.
    fft1d_prefetch = 64u,
};

/// Largest radix a stage computes in registers; a larger (odd) radix runs as a direct DFT stage
/// whose butterfly is spread over several work-items (inputs in LDS, O(R^2) symmetric-pair
/// sums, the role of the reference's basic_esum fallback, src/base/mixed_radix_fft.cpp:46-53).
constexpr int fft1d_max_register_radix = 64;
inline bool is_direct_radix(int r) { return r > fft1d_max_register_radix; }

/// Number of half-precision terms the DFT matrix is split into on the MFMA stage.
inline unsigned mfma_split_terms(unsigned flags) { return (flags & fft1d_input_bf16) ? 3u : 2u; }

/// Optional overrides of the tiling heuristics (0 / empty = choose automatically).
struct BBFFT_EXPORT fft1d_tuning {
    std::vector<int> factorization;
    std::size_t Nt = 0;
    std::size_t Mb = 0;
    std::size_t Kb = 0;
    std::size_t lds_pad = 0;
    bool no_pad = false;
    unsigned flags = 0;
    bool has_flags = false; ///< use `flags` instead of the default
    /// Waves per SIMD the register allocator must allow (0 = its default); a higher target
    /// trades registers (possibly spills) for occupancy.
    unsigned waves_per_eu = 0;
    /// Persistent launch: this many work-groups per CU loop over the tiles (0 = one work-group
    /// per tile). Does not change the kernel, only its grid.
    unsigned groups_per_cu = 0;
};

/// Choose a kernel configuration for a 1-D configuration (dim == 1). `variant` (fft1d_flag
/// bits such as fft1d_adjoint) is added to the flags whatever the tuning says.
BBFFT_EXPORT fft1d_configuration configure_fft1d(configuration const &cfg,
                                                 device_info const &info,
                                                 fft1d_tuning const &tuning = {},
                                                 unsigned variant = 0u);
/**
 * @brief Emit HIP source (no includes; hiprtc provides the HIP device builtins) for one kernel.
 *
 * When `callbacks` is given its source is placed in front of the kernel; OpenCL-C callbacks
 * (the reference's language) get the address-space qualifiers mapped away and are compiled as
 * host+device functions.
 */
BBFFT_EXPORT void generate_fft1d(std::ostream &os, fft1d_configuration const &cfg,
                                 std::string_view name = {},
                                 user_module const *callbacks = nullptr);

/**
 * @brief Two consecutive passes of a 2-D/3-D plan fused into one kernel over "planes".
 *
 * The intermediate tensor between pass a and pass b is packed, so it splits into contiguous
 * planes of `plane_elems` complex elements, each holding ka K-columns of pass a's output and kb
 * K-columns of pass b's input. Work-group f computes pass a for its plane (one tile: all
 * M-columns and all ka K-columns) into LDS, then pass b from LDS to global memory; the plane
 * shares the LDS with the passes' stage scratch (barriers separate the uses). HBM traffic is
 * one read of the input and one write of the output instead of two round trips.
 */
struct BBFFT_EXPORT fused2_configuration {
    fft1d_configuration a;          ///< first pass (reads global memory, writes the plane)
    fft1d_configuration b;          ///< second pass (reads the plane, writes global memory)
    std::size_t ka;                 ///< K-columns of pass a per plane
    std::size_t kb;                 ///< K-columns of pass b per plane
    std::size_t planes;             ///< number of planes
    std::size_t plane_elems;        ///< complex elements per plane
    std::size_t in_plane_stride;    ///< input elements between planes (ka * a.istride[2])
    std::size_t out_plane_stride;   ///< output elements between planes (kb * b.ostride[2])
    std::size_t group_size;         ///< work-items per work-group
    /// 32-bit words of one plane's input (rounded up to 64) that each work-group streams into
    /// an LDS staging buffer with direct global-to-LDS loads while it still works on the
    /// previous plane (persistent grid); 0: pass a reads global memory itself.
    std::size_t stage_words = 0;
    /// Bytes per work-item of each direct global-to-LDS staging load: 16 (whole 1 KB blocks
    /// per wave) or 4. 16-byte loads need a 16-byte aligned input, which the kernel checks;
    /// otherwise it takes the 4-byte loads.
    unsigned stage_load_bytes = 4;
    /// The staging loads carry the non-temporal cache policy (the input is read once).
    bool stage_load_nt = false;
    /// Barriers order LDS accesses only: each work-item's global accesses within a plane are
    /// its own, so waits for outstanding global loads and stores are not needed at barriers.
    bool lds_barriers = false;
    /// Twiddle tables copied into LDS once per work-group (no global loads inside the plane
    /// loop besides the staging loads).
    bool lds_tables = false;
    /// With lds_tables: the LDS copies of the stage twiddle tables keep their rows (one per
    /// j mod p) R + 1 entries apart, so work-items reading different rows at once do not
    /// share LDS banks.
    bool lds_table_pad = false;
    /// Occupancy target of the register allocator (waves per SIMD, 0: the compiler's default);
    /// 8 lets two 1024-work-item groups share a CU when their LDS fits.
    unsigned waves_per_eu = 0;
    /// First wave that issues the staging loads (0: all waves). With LDS-only barriers and LDS
    /// tables, waves from this one on have no work in pass b and issue no global stores, so
    /// waiting for their staging loads does not wait for the plane's stores (which share the
    /// wait counter with the loads).
    unsigned copy_wave0 = 0;
    /// The next plane's staging loads start as soon as pass a's first stage has read the
    /// staged input (instead of after pass a).
    bool early_copy = false;
    /// Staged input rows whose length is a multiple of 512 bytes alias in the LDS banks when a
    /// wave reads one segment of each of several rows at once. With stage_swizzle_rows = R
    /// (a power of two, 0 = off), the 16-byte chunks of staged row r are permuted: segment s
    /// (stage_swizzle_elems complex elements) of the row holds the row's segment
    /// s ^ (r % R), so the rows read together land in different banks. Pass a applies the same
    /// XOR to its input element indices.
    unsigned stage_swizzle_rows = 0;
    unsigned stage_swizzle_elems = 0;

    std::string identifier() const;
    std::size_t lds_bytes() const; ///< plane and stage scratch (without the staging buffer)
};

/**
 * @brief Configure the fused kernel for 1-D passes pa then pb of one plan (as decompose makes
 * them); returns false when they cannot be fused on this device (plane or stage scratch over
 * the LDS size, or a pass that does not fit one work-group).
 */
BBFFT_EXPORT bool configure_fused2(configuration const &pa, configuration const &pb,
                                   device_info const &info, unsigned variant_a,
                                   unsigned variant_b, fused2_configuration &out);
/// Emit the HIP source of the fused kernel `name(in, out, planes)`.
BBFFT_EXPORT void generate_fused2(std::ostream &os, fused2_configuration const &fc,
                                  std::string_view name = {});

/**
 * @brief Persistent kernel of a ragged batch.
 *
 * Every distinct (N, M, direction) of the jobs is a variant: an fft1d_configuration whose
 * factorisation, work-items per transform and M tile come from configure_fft1d (tuning table or
 * heuristics) and whose K tile is widened so that Mb * Nt * Kb fills the common work-group of
 * `group_size` work-items. The tiles of all jobs form one linear tile space; work-group g
 * processes tiles g, g + grid, ... and finds each tile's job by binary search in the job table.
 */
struct BBFFT_EXPORT ragged_kernel_configuration {
    precision fp;
    std::size_t group_size;                    ///< work-items per work-group
    std::vector<fft1d_configuration> variants; ///< one per distinct (N, M, direction)
    std::vector<unsigned> job_variant;         ///< variant of each job
    std::vector<std::size_t> order;            ///< job index at each table position
    std::vector<std::uint64_t> tile_begin;     ///< first tile of each table position, + total
    unsigned waves_per_eu = 0;                 ///< occupancy target of the kernel (0: none)

    std::uint64_t total_tiles() const { return tile_begin.empty() ? 0 : tile_begin.back(); }
    /// Kernel symbol name; depends on the variants, the work-group size and the occupancy
    /// target only, so the compiled module serves any job list with the same variants.
    std::string identifier() const;
    /// Bytes of LDS (the largest of the variants).
    std::size_t lds_bytes() const;
};

/// Configure the persistent kernel for a ragged batch (validates the jobs).
BBFFT_EXPORT ragged_kernel_configuration configure_ragged(ragged_configuration const &cfg,
                                                          device_info const &info);
/// Emit the HIP source of the persistent kernel `name(in, out, table, njobs)`.
BBFFT_EXPORT void generate_ragged(std::ostream &os, ragged_kernel_configuration const &rk,
                                  std::string_view name = {});
/**
 * @brief Device job table for a launch of `grid` work-groups: tile_begin[0..njobs] (njobs + 1
 * entries, the last is the total), njobs records {in_offset, out_offset, K, variant} in table
 * order (offsets in complex elements), then the work split: first tile of each work-group
 * (grid + 1 entries, the last is the total) and the table position of the job holding that
 * tile (grid entries). Work-group g processes the contiguous tiles [first[g], first[g + 1]);
 * the boundaries balance the elements processed (tile count times elements per tile).
 */
BBFFT_EXPORT std::vector<std::uint64_t> ragged_job_table(ragged_configuration const &cfg,
                                                         ragged_kernel_configuration const &rk,
                                                         std::size_t grid);

} // namespace bbfft

#endif // BBFFT_DETAIL_GENERATOR_IMPL_HPP
