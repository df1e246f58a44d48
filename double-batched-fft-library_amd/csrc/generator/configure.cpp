// Tiling and factorisation heuristics for the gfx950 kernels.
// Plays the role of the reference's configure_small_batch_fft / configure_factor2_slm_fft
// (src/base/generator/small_batch_fft.cpp:15-58, factor2_slm_fft.cpp:16-86), re-derived for
// wave64, 160 KiB LDS and Stockham stages exchanged through LDS.
#include "../base/math.hpp"
#include "bbfft/bad_configuration.hpp"
#include "bbfft/detail/generator_impl.hpp"
#include "heuristics.hpp"
#include "tuning_table.hpp"

#include <algorithm>
#include <cstdlib>
#include <numeric>
#include <sstream>
#include <string>

namespace bbfft {

std::size_t fft1d_configuration::N_fft() const {
    bool real = type != transform_type::c2c;
    return real && N % 2 == 0 ? N / 2 : N;
}

bool fft1d_configuration::pairs_k() const { return type != transform_type::c2c && N % 2 == 1; }

std::size_t fft1d_configuration::K_columns(std::size_t K) const {
    return pairs_k() ? (K + 1) / 2 : K;
}

namespace {
std::size_t padded(std::size_t n, std::size_t pad) { return pad ? n + n / pad : n; }

bool needs_lds(fft1d_configuration const &c) {
    return c.factorization.size() > 1 || c.type != transform_type::c2c || c.direct_radix();
}
} // namespace

int fft1d_configuration::direct_radix() const {
    for (auto f : factorization) {
        if (is_direct_radix(f)) {
            return f;
        }
    }
    return 0;
}

std::size_t fft1d_configuration::lds_bytes() const {
    if (!needs_lds(*this)) {
        return 0;
    }
    std::size_t np = padded(N_fft() - 1, lds_pad) + 1;
    // the direct stage keeps its table of roots of unity in LDS
    std::size_t const table = static_cast<std::size_t>(direct_radix());
    return (Mb * Kb * np + table) * 2 * static_cast<std::size_t>(fp);
}

std::string fft1d_configuration::identifier() const {
    std::ostringstream oss;
    oss << "bbfft_" << (direction < 0 ? 'm' : 'p') << (direction < 0 ? -direction : direction)
        << "_M" << M << "_Mb" << Mb << "_N" << N << "_f";
    for (std::size_t i = 0; i < factorization.size(); ++i) {
        oss << (i ? "x" : "") << factorization[i];
    }
    oss << "_Nt" << Nt << "_Kb" << Kb << "_f" << static_cast<int>(fp) * 8 << '_'
        << to_string(type) << "_is" << istride[0] << '_' << istride[1] << '_' << istride[2]
        << "_os" << ostride[0] << '_' << ostride[1] << '_' << ostride[2] << "_p" << lds_pad
        << "_in" << inplace_unsupported;
    if (flags) {
        oss << "_g" << flags;
    }
    if (waves_per_eu) {
        oss << "_w" << waves_per_eu;
    }
    if (load_function) {
        oss << '_' << load_function;
    }
    if (store_function) {
        oss << '_' << store_function;
    }
    if (callback_hash) {
        oss << "_cb" << std::hex << callback_hash;
    }
    return oss.str();
}

namespace gen {

std::uint64_t callback_hash(user_module const &um) {
    if (!um) {
        return 0;
    }
    std::uint64_t h = 0xcbf29ce484222325ull;
    auto mix = [&h](unsigned char b) {
        h ^= b;
        h *= 0x100000001b3ull;
    };
    for (std::size_t i = 0; i < um.length; ++i) {
        mix(static_cast<unsigned char>(um.data[i]));
    }
    mix(static_cast<unsigned char>(um.language));
    return h ? h : 1;
}

/// Stage radices for a complex FFT of length n: one stage up to `single_max`, otherwise the
/// fewest stages whose largest radix does not exceed `radix_max` (factors closest to n^(1/L)).
std::vector<int> default_factorization(std::size_t n, std::size_t single_max,
                                       std::size_t radix_max) {
    if (n <= single_max) {
        return {static_cast<int>(n)};
    }
    for (unsigned L = 2; L <= 4; ++L) {
        auto f = factor(static_cast<unsigned>(n), L);
        // drop unit factors
        std::vector<int> g;
        for (auto v : f) {
            if (v > 1) {
                g.push_back(static_cast<int>(v));
            }
        }
        int fmax = *std::max_element(g.begin(), g.end());
        if (static_cast<std::size_t>(fmax) <= radix_max || L == 4) {
            // Largest radix first: the first stage then has the fewest butterflies, which sets
            // the work-item count, and later stages loop over more butterflies per work-item.
            std::sort(g.begin(), g.end(), std::greater<int>());
            return g;
        }
    }
    return {static_cast<int>(n)};
}

} // namespace gen

fft1d_configuration configure_fft1d(configuration const &cfg, device_info const &info,
                                    fft1d_tuning const &tuning_in, unsigned variant) {
    if (cfg.dim != 1) {
        throw bad_configuration("configure_fft1d requires a 1-D configuration");
    }
    if (cfg.istride[0] != 1 || cfg.ostride[0] != 1) {
        throw bad_configuration("Stride s_0 != 1 is not supported");
    }
    if ((cfg.type == transform_type::r2c && cfg.dir != direction::forward) ||
        (cfg.type == transform_type::c2r && cfg.dir != direction::backward)) {
        throw bad_configuration("r2c must be forward and c2r must be backward");
    }
    std::size_t const M = cfg.shape[0], N = cfg.shape[1], K = cfg.shape[2];
    if (M == 0 || N == 0 || K == 0) {
        throw bad_configuration("Tensor shape must not contain zeros");
    }
    // Without explicit tuning, a measured entry of the persisted table (if any) is used; its
    // K tile is clamped to the number of K columns below.
    fft1d_tuning tuning = tuning_in;
    bool const explicit_tuning = !tuning_in.factorization.empty() || tuning_in.Nt ||
                                 tuning_in.Mb || tuning_in.Kb || tuning_in.lds_pad ||
                                 tuning_in.no_pad || tuning_in.has_flags;
    unsigned const all_flags = (tuning_in.has_flags ? tuning_in.flags : 0u) | variant;
    bool const half_in = (all_flags & (fft1d_input_f16 | fft1d_input_bf16)) != 0;
    bool const mfma = (all_flags & fft1d_mfma) != 0;
    if (!explicit_tuning && info.type == device_type::gpu && info.max_subgroup_size() == 64) {
        gen::lookup_tuning(cfg.fp, cfg.type, N, M, all_flags, tuning);
        if (char const *m = std::getenv("BBFFT_TABLE_FLAGS_CLEAR")) { // A/B: drop entry flags
            tuning.flags &= ~static_cast<unsigned>(std::strtoul(m, nullptr, 0));
        }
        // The table is measured on plain kernels; the load-first tile loop needs direct global
        // loads, so a plan with callbacks keeps the entry's tiling without it, and without the
        // entry's occupancy target and persistent grid: the callback's code adds registers (a
        // filter-multiplying load callback on the 128-point pass of 32 x 128^3 spilled 46
        // VGPRs at 4 waves per SIMD and needs 146 without the cap, no spills).
        if (cfg.callbacks) {
            tuning.flags &= ~static_cast<unsigned>(fft1d_prefetch);
            tuning.waves_per_eu = 0;
            tuning.groups_per_cu = 0;
            if (char const *w = std::getenv("BBFFT_CALLBACK_WAVES")) { // A/B
                tuning.waves_per_eu = static_cast<unsigned>(std::atoi(w));
            }
            // A kernel with a load callback streams its stores past the last-level cache: the
            // callback may read side data (a filter read once per batch) that the kernel's
            // output would otherwise evict. BBFFT_CALLBACK_NT=off keeps the entry's stores.
            char const *nt = std::getenv("BBFFT_CALLBACK_NT");
            if (cfg.callbacks.load_function && tuning.has_flags &&
                !(nt && std::string(nt) == "off")) {
                tuning.flags |= fft1d_nontemporal_store;
            }
        }
    }

    fft1d_configuration c = {};
    c.direction = static_cast<int>(cfg.dir);
    c.M = M;
    c.N = N;
    c.fp = cfg.fp;
    c.type = cfg.type;
    c.istride = {cfg.istride[0], cfg.istride[1], cfg.istride[2]};
    c.ostride = {cfg.ostride[0], cfg.ostride[1], cfg.ostride[2]};
    c.load_function = cfg.callbacks.load_function;
    c.store_function = cfg.callbacks.store_function;
    c.callback_hash = gen::callback_hash(cfg.callbacks);
    std::size_t const nfft = c.N_fft();

    if (half_in) {
        if ((all_flags & fft1d_input_f16) && (all_flags & fft1d_input_bf16)) {
            throw bad_configuration("Choose one half-precision input format");
        }
        if (cfg.type != transform_type::c2c || cfg.fp != precision::f32) {
            throw bad_configuration("Half-precision input requires a c2c fp32 transform");
        }
        if (cfg.callbacks) {
            throw bad_configuration("Half-precision input does not support callbacks");
        }
    }
    if (mfma) {
        if (!half_in) {
            throw bad_configuration("The MFMA stage requires half-precision input");
        }
        if (nfft % 16 != 0 || nfft < 32) {
            throw bad_configuration("The MFMA stage requires N divisible by 16 and N >= 32");
        }
        if (!tuning.factorization.empty() && tuning.factorization[0] != 16) {
            throw bad_configuration("The MFMA stage requires a first radix of 16");
        }
    }
    if (!tuning.factorization.empty()) {
        std::size_t p = 1;
        for (auto f : tuning.factorization) {
            if (f < 1) {
                throw bad_configuration("factorization entries must be positive");
            }
            p *= static_cast<std::size_t>(f);
        }
        if (p != nfft) {
            throw bad_configuration("factorization does not multiply to the FFT length");
        }
        c.factorization = tuning.factorization;
    } else {
        // Radices up to 16: a larger radix means more elements per work-item (registers) and
        // fewer work-items per transform. On MI355X the 2-stage radix-20..32 splits of N = 243,
        // 384, 500, 512 at M <= 4 reached 3.1-5.8 TB/s against 5.7-6.6 TB/s with radix <= 16.
        std::size_t const single_max = 16;
        std::size_t const radix_max = 16;
        c.factorization = gen::default_factorization(nfft, single_max, radix_max);
        if (mfma) {
            c.factorization = gen::default_factorization(nfft / 16, 16, 16);
            c.factorization.insert(c.factorization.begin(), 16);
        }
    }
    // Radices above the register limit run as a direct DFT stage (one per kernel, odd radix:
    // the symmetric-pair sums need no self-paired index), as the reference falls back to
    // basic_esum for prime factors above its register budget.
    int direct = 0;
    for (auto f : c.factorization) {
        if (is_direct_radix(f)) {
            if (f % 2 == 0) {
                throw bad_configuration("Radix " + std::to_string(f) +
                                        " exceeds the in-register limit (" +
                                        std::to_string(fft1d_max_register_radix) +
                                        ") and is even; split it into smaller radices");
            }
            if (direct) {
                throw bad_configuration("N = " + std::to_string(N) +
                                        " has two factors above the in-register limit (" +
                                        std::to_string(fft1d_max_register_radix) +
                                        "); at most one is supported");
            }
            direct = f;
        }
    }
    if (direct && mfma) {
        throw bad_configuration("The MFMA stage does not combine with a direct DFT stage");
    }
    // The MFMA stage does not use the work-items-per-transform split; later stages set it.
    int const rmax = mfma ? *std::max_element(c.factorization.begin() + 1, c.factorization.end())
                          : *std::max_element(c.factorization.begin(), c.factorization.end());
    if (direct) {
        // Work-item t of a transform serves butterfly t % (N_fft / R) of the direct stage and
        // output pairs t / (N_fft / R) + b Tk of it; Tk is chosen so that each work-item keeps
        // about 8 (fp32) or 6 (fp64) output pairs in registers.
        std::size_t const NRd = nfft / static_cast<std::size_t>(direct);
        std::size_t const units = static_cast<std::size_t>(direct - 1) / 2 + 1;
        std::size_t const per_item = cfg.fp == precision::f32 ? 8 : 6;
        std::size_t const Tk = (units + per_item - 1) / per_item;
        c.Nt = tuning.Nt ? tuning.Nt : NRd * Tk;
        if (c.Nt % NRd != 0) {
            throw bad_configuration("With a direct DFT stage of radix " + std::to_string(direct) +
                                    " the work-items per transform must be a multiple of " +
                                    std::to_string(NRd));
        }
    } else {
        c.Nt = tuning.Nt ? tuning.Nt : nfft / static_cast<std::size_t>(rmax);
    }

    std::size_t const wg_max = std::min<std::size_t>(1024, info.max_work_group_size);
    if (c.Nt == 0 || c.Nt > wg_max) {
        throw bad_configuration("Invalid number of work-items per transform");
    }
    if (tuning.Mb) {
        c.Mb = std::min(tuning.Mb, M);
    } else {
        std::size_t mmax = std::max<std::size_t>(1, std::min<std::size_t>(64, 512 / c.Nt));
        std::size_t mg = (M + mmax - 1) / mmax;
        c.Mb = (M + mg - 1) / mg;
    }
    if (c.Mb * c.Nt > wg_max) {
        throw bad_configuration("M tile times work-items per transform exceeds work-group size");
    }

    c.lds_pad = 0;
    if (tuning.lds_pad) {
        c.lds_pad = tuning.lds_pad;
    } else if (!tuning.no_pad && (c.factorization.size() > 1 || cfg.type != transform_type::c2c)) {
        c.lds_pad = nfft >= 32 ? 16 : 0;
    }
    if (c.lds_pad & (c.lds_pad - 1)) {
        throw bad_configuration("lds_pad must be a power of two");
    }

    std::size_t const kcols = c.K_columns(K);
    std::size_t const lds_limit = std::min<std::size_t>(info.local_memory_size, 64 * 1024);
    if (!tuning.Mb) {
        // Shrink the M tile until one K-column tile fits the LDS budget.
        c.Kb = 1;
        std::size_t mg = (M + c.Mb - 1) / c.Mb;
        while (c.Mb > 1 && c.lds_bytes() > lds_limit) {
            ++mg;
            c.Mb = (M + mg - 1) / mg;
        }
    }
    if (tuning.Kb) {
        c.Kb = std::min(tuning.Kb, kcols);
    } else {
        std::size_t target = 256;
        std::size_t kb = std::max<std::size_t>(1, target / (c.Mb * c.Nt));
        kb = std::min(kb, kcols);
        c.Kb = kb;
        while (c.Kb > 1 && c.lds_bytes() > lds_limit) {
            --c.Kb;
        }
    }
    if (mfma) {
        // Matrix instructions read all 64 lanes of a wave: whole waves only (the extra K
        // columns are inactive when K is smaller).
        std::size_t const step = 64 / std::gcd<std::size_t>(c.Mb * c.Nt, 64);
        c.Kb = (c.Kb + step - 1) / step * step;
        if (c.lds_bytes() > lds_limit) {
            throw bad_configuration("MFMA kernel tile exceeds the LDS budget");
        }
    }
    if (c.work_group_size() > wg_max) {
        throw bad_configuration("Work-group size exceeds the device limit");
    }
    if (c.lds_bytes() > info.local_memory_size) {
        throw bad_configuration("Work-group exceeds the local memory size");
    }
    // Access hints without a measured entry: streaming (non-temporal) loads and stores when a
    // work-group's M run is at least 24 contiguous bytes. Measured on MI355X for untuned shapes
    // (heuristic tiling, about 1 GB per buffer, flags 0 vs 3): fp32 M=4 N=500 +11 %, M=3
    // N=384 +8 %, M=4 N=128 +2 %, M=16 N=64 +4 %, fp64 M=2 N=500 +5 %; fp32 M=1 N=60 -55 %,
    // M=1 N=100 -16 %, M=2 N=243 -13 %, r2c M=1 N=500 -18 %, c2r M=2 N=243 -6 %.
    unsigned default_hints = 0;
    if (info.type == device_type::gpu &&
        c.Mb * 2 * static_cast<std::size_t>(c.fp) >= 24) {
        default_hints = fft1d_nontemporal_load | fft1d_nontemporal_store;
    }
    c.flags = (tuning.has_flags ? tuning.flags : default_hints) | variant;
    if (tuning.waves_per_eu > 8) {
        throw bad_configuration("waves_per_eu must be 0..8");
    }
    c.waves_per_eu = tuning.waves_per_eu;
    c.groups_per_cu = tuning.groups_per_cu;
    if (c.flags & fft1d_prefetch) {
        if (cfg.type == transform_type::c2r || cfg.callbacks || (c.flags & fft1d_mfma) ||
            direct) {
            throw bad_configuration("Tile prefetching requires a c2c or r2c kernel without "
                                    "callbacks, MFMA or direct DFT stage");
        }
    }
    if ((c.flags & fft1d_adjoint) && cfg.type == transform_type::c2c) {
        throw bad_configuration("The adjoint variant applies to real transforms only");
    }
    c.inplace_unsupported = cfg.type != transform_type::c2c && c.Mb < M;
    return c;
}

} // namespace bbfft
