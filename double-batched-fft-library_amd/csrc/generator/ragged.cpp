// Ragged batches: variant selection, job table and the persistent kernel.
//
// No counterpart in the reference (one kernel per shape there). The per-variant tile code is
// the same generator output as the plain kernels (emitted as device functions), so a job's
// arithmetic is the arithmetic of the plain plan with the same factorisation.
#include "bbfft/bad_configuration.hpp"
#include "bbfft/detail/generator_impl.hpp"
#include "emitter.hpp"
#include "heuristics.hpp"
#include "tuning_table.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdint>
#include <iomanip>
#include <map>
#include <numeric>
#include <ostream>
#include <sstream>
#include <tuple>

namespace bbfft {

namespace {

/// Work-items per work-group of the persistent kernel (at least this many).
constexpr std::size_t ragged_min_group = 256;
/// LDS budget of one variant: 160 KB / 5, so that the five work-groups per CU that the
/// occupancy target (5 waves per SIMD, 4 waves per group) allows also fit in LDS.
constexpr std::size_t ragged_lds_limit = 32 * 1024;
/// Largest radix of a variant. All variants share one kernel, whose register allocation is the
/// largest of the variants', so one wide radix would lower the occupancy of every job.
constexpr int ragged_radix_max = 12;
/// K used to configure a variant: large enough that the K tile is never clamped by K.
constexpr std::size_t variant_K = std::size_t(1) << 30;

configuration variant_configuration(precision fp, std::size_t M, std::size_t N, direction dir) {
    configuration c{1, {M, N, variant_K}, fp, dir, transform_type::c2c};
    c.set_strides_default(false);
    return c;
}

std::uint64_t fnv1a(std::string const &s) {
    std::uint64_t h = 1469598103934665603ull;
    for (unsigned char ch : s) {
        h ^= ch;
        h *= 1099511628211ull;
    }
    return h;
}

std::string variant_name(std::string const &kernel, std::size_t i) {
    return kernel + "_v" + std::to_string(i);
}

} // namespace

std::string ragged_kernel_configuration::identifier() const {
    std::string all;
    for (auto const &v : variants) {
        all += v.identifier();
        all += ';';
    }
    all += "w" + std::to_string(waves_per_eu);
    std::ostringstream oss;
    oss << "bbfft_ragged_f" << static_cast<int>(fp) * 8 << "_G" << group_size << "_V"
        << variants.size() << "_" << std::hex << std::setw(16) << std::setfill('0') << fnv1a(all);
    return oss.str();
}

std::size_t ragged_kernel_configuration::lds_bytes() const {
    std::size_t b = 0;
    for (auto const &v : variants) {
        b = std::max(b, v.lds_bytes());
    }
    return b;
}

ragged_kernel_configuration configure_ragged(ragged_configuration const &cfg,
                                             device_info const &info) {
    if (cfg.fp != precision::f32 && cfg.fp != precision::f64) {
        throw bad_configuration("Unsupported precision");
    }
    if (cfg.jobs.empty()) {
        throw bad_configuration("A ragged batch needs at least one job");
    }
    ragged_kernel_configuration rk = {};
    rk.fp = cfg.fp;
    // distinct (N, M, direction), in sorted order so that the kernel depends on the set only
    using key_t = std::tuple<std::size_t, std::size_t, int>;
    std::map<key_t, unsigned> index;
    for (std::size_t j = 0; j < cfg.jobs.size(); ++j) {
        auto const &job = cfg.jobs[j];
        if (job.M == 0 || job.N == 0 || job.K == 0) {
            throw bad_configuration("Ragged job " + std::to_string(j) +
                                    ": shape must not contain zeros");
        }
        if (job.dir != direction::forward && job.dir != direction::backward) {
            throw bad_configuration("Ragged job " + std::to_string(j) + ": invalid direction");
        }
        index.emplace(key_t{job.N, job.M, static_cast<int>(job.dir)}, 0u);
    }
    std::vector<key_t> keys;
    for (auto &[k, i] : index) {
        i = static_cast<unsigned>(keys.size());
        keys.push_back(k);
    }
    rk.job_variant.resize(cfg.jobs.size());
    for (std::size_t j = 0; j < cfg.jobs.size(); ++j) {
        auto const &job = cfg.jobs[j];
        rk.job_variant[j] = index.at(key_t{job.N, job.M, static_cast<int>(job.dir)});
    }
    // Factorisation, work-items per transform, M tile, LDS padding and access hints: the
    // table's ragged entry for (N, M) (tuned under this kernel's constraints), else the
    // plain-kernel configuration (tuning table or heuristics) with the radix capped.
    std::vector<fft1d_tuning> tune(keys.size());
    // BBFFT_RAGGED_GROUP / BBFFT_RAGGED_LDS override the common group size and the LDS budget
    // of a variant (measurements)
    std::size_t G = ragged_min_group;
    std::size_t lds_limit = ragged_lds_limit;
    if (char const *g = std::getenv("BBFFT_RAGGED_GROUP")) {
        G = std::max<std::size_t>(64, static_cast<std::size_t>(std::atoi(g)));
    }
    if (char const *l = std::getenv("BBFFT_RAGGED_LDS")) {
        lds_limit = static_cast<std::size_t>(std::atoi(l));
    }
    for (std::size_t v = 0; v < keys.size(); ++v) {
        auto const &[N, M, d] = keys[v];
        fft1d_tuning &t = tune[v];
        fft1d_tuning rt;
        if (gen::lookup_tuning(cfg.fp, transform_type::c2c, N, M, gen::tuning_ragged, rt) &&
            !rt.factorization.empty() && rt.Mb * rt.Nt <= std::max(G, ragged_min_group)) {
            t.factorization = rt.factorization;
            t.Nt = rt.Nt;
            t.Mb = rt.Mb;
            t.lds_pad = rt.lds_pad;
            t.no_pad = rt.lds_pad == 0;
            t.flags = rt.flags;
            t.has_flags = true;
            G = std::max(G, t.Mb * t.Nt);
            continue;
        }
        auto b = configure_fft1d(variant_configuration(cfg.fp, M, N, static_cast<direction>(d)),
                                 info);
        t.factorization = b.factorization;
        t.Nt = b.Nt;
        int const rmax = *std::max_element(b.factorization.begin(), b.factorization.end());
        if (rmax > ragged_radix_max) {
            std::size_t const nfft = b.N_fft();
            t.factorization = gen::default_factorization(nfft, ragged_radix_max, ragged_radix_max);
            t.Nt = nfft / static_cast<std::size_t>(
                              *std::max_element(t.factorization.begin(), t.factorization.end()));
        }
        t.Mb = b.Mb;
        t.lds_pad = b.lds_pad;
        t.no_pad = b.lds_pad == 0;
        t.flags = b.flags;
        t.has_flags = true;
        G = std::max(G, t.Mb * t.Nt);
    }
    G = std::min<std::size_t>(G, std::min<std::size_t>(1024, info.max_work_group_size));
    // K tile: fill the common work-group, within the LDS budget.
    for (std::size_t v = 0; v < keys.size(); ++v) {
        auto const &[N, M, d] = keys[v];
        fft1d_tuning &t = tune[v];
        t.Kb = std::max<std::size_t>(1, G / (t.Mb * t.Nt));
        fft1d_configuration c;
        for (;;) {
            c = configure_fft1d(variant_configuration(cfg.fp, M, N, static_cast<direction>(d)),
                                info, t);
            if (c.lds_bytes() <= lds_limit || t.Kb == 1) {
                break;
            }
            t.Kb = t.Kb / 2;
        }
        if (c.work_group_size() > G) {
            throw bad_configuration("Ragged variant exceeds the common work-group size");
        }
        if (c.direct_radix()) {
            throw bad_configuration("Ragged job length " + std::to_string(N) +
                                    " needs a direct DFT stage (a prime factor above " +
                                    std::to_string(fft1d_max_register_radix) +
                                    "); ragged batches support register radices only");
        }
        rk.variants.push_back(c);
    }
    rk.group_size = G;
    // occupancy target (BBFFT_RAGGED_WAVES overrides it for measurements)
    rk.waves_per_eu = cfg.fp == precision::f32 ? 5u : 0u;
    if (char const *w = std::getenv("BBFFT_RAGGED_WAVES")) {
        rk.waves_per_eu = static_cast<unsigned>(std::atoi(w));
    }
    // Table order: grouped by variant (work-groups running at the same time then tend to run
    // the same code), jobs of one variant in their given order.
    rk.order.resize(cfg.jobs.size());
    std::iota(rk.order.begin(), rk.order.end(), std::size_t(0));
    std::stable_sort(rk.order.begin(), rk.order.end(), [&](std::size_t a, std::size_t b) {
        return rk.job_variant[a] < rk.job_variant[b];
    });
    rk.tile_begin.resize(cfg.jobs.size() + 1);
    std::uint64_t tiles = 0;
    for (std::size_t p = 0; p < rk.order.size(); ++p) {
        auto const &job = cfg.jobs[rk.order[p]];
        auto const &c = rk.variants[rk.job_variant[rk.order[p]]];
        rk.tile_begin[p] = tiles;
        std::uint64_t mg = (job.M + c.Mb - 1) / c.Mb;
        tiles += mg * ((job.K + c.Kb - 1) / c.Kb);
    }
    rk.tile_begin.back() = tiles;
    return rk;
}

std::vector<std::uint64_t> ragged_job_table(ragged_configuration const &cfg,
                                            ragged_kernel_configuration const &rk,
                                            std::size_t grid) {
    std::size_t const n = cfg.jobs.size();
    if (rk.order.size() != n || rk.tile_begin.size() != n + 1) {
        throw bad_configuration("Ragged kernel configuration does not match the jobs");
    }
    if (grid == 0) {
        throw bad_configuration("The work split needs at least one work-group");
    }
    std::vector<std::uint64_t> table(n + 1 + 4 * n + (grid + 1) + grid);
    std::copy(rk.tile_begin.begin(), rk.tile_begin.end(), table.begin());
    std::vector<double> weight(n); // elements per tile of each table position
    for (std::size_t p = 0; p < n; ++p) {
        auto const &job = cfg.jobs[rk.order[p]];
        auto const &v = rk.variants[rk.job_variant[rk.order[p]]];
        std::uint64_t *r = table.data() + n + 1 + 4 * p;
        r[0] = job.in_offset;
        r[1] = job.out_offset;
        r[2] = job.K;
        r[3] = rk.job_variant[rk.order[p]];
        weight[p] = static_cast<double>(v.Mb * v.Kb * v.N);
    }
    // Work split: boundary g sits at g / grid of the total weight.
    std::vector<double> cum(n + 1, 0.0);
    for (std::size_t p = 0; p < n; ++p) {
        cum[p + 1] = cum[p] + weight[p] * static_cast<double>(rk.tile_begin[p + 1] -
                                                              rk.tile_begin[p]);
    }
    std::uint64_t *first = table.data() + 5 * n + 1;
    std::uint64_t *pos = first + grid + 1;
    std::size_t p = 0;
    for (std::size_t g = 0; g <= grid; ++g) {
        std::uint64_t t = rk.total_tiles();
        if (g < grid) {
            double const target = cum[n] * static_cast<double>(g) / static_cast<double>(grid);
            while (p + 1 < n && cum[p + 1] <= target) {
                ++p;
            }
            double const k = std::ceil((target - cum[p]) / weight[p]);
            t = std::min<std::uint64_t>(rk.tile_begin[p] + static_cast<std::uint64_t>(
                                                               std::max(0.0, k)),
                                        rk.tile_begin[p + 1]);
            t = std::max<std::uint64_t>(t, g ? first[g - 1] : 0);
        }
        first[g] = t;
    }
    // job holding each first tile: the last position whose first tile is <= it
    for (std::size_t g = 0; g < grid; ++g) {
        auto it = std::upper_bound(rk.tile_begin.begin(), rk.tile_begin.end() - 1, first[g]);
        pos[g] = static_cast<std::uint64_t>(std::max<std::ptrdiff_t>(
            0, (it - rk.tile_begin.begin()) - 1));
    }
    return table;
}

void generate_ragged(std::ostream &os, ragged_kernel_configuration const &rk,
                     std::string_view name) {
    if (rk.variants.empty()) {
        throw bad_configuration("A ragged kernel needs at least one variant");
    }
    std::string const n = name.empty() ? rk.identifier() : std::string(name);
    std::string const T2 = rk.fp == precision::f32 ? "float2" : "double2";
    os << hip_prelude();
    for (std::size_t v = 0; v < rk.variants.size(); ++v) {
        generate_fft1d_body(os, rk.variants[v], variant_name(n, v), rk.group_size);
    }
    std::size_t const lds_elems = std::max<std::size_t>(1, rk.lds_bytes() / (2 * static_cast<std::size_t>(rk.fp)));
    // One kernel holds every variant, so its register allocation follows the heaviest one;
    // for fp32 variants (radix <= 12) the allocator reaches 5 waves per SIMD (96 VGPRs)
    // without spilling when asked to (measured on MI355X for the 4096-job serving mix: 78.4 ms
    // against 85.4 ms with radix <= 16 at 4 waves).
    std::string const occupancy =
        rk.waves_per_eu ? " __attribute__((amdgpu_waves_per_eu(" +
                              std::to_string(rk.waves_per_eu) + ")))"
                        : "";
    os << "extern \"C\" __global__ void __launch_bounds__(" << rk.group_size << ")"
       << occupancy << " " << n
       << "(const " << T2 << "* in, " << T2
       << "* out, const unsigned long* table, unsigned long njobs) {\n"
       << "  __shared__ " << T2 << " lds_buf[" << lds_elems << "];\n"
       << "  const unsigned long* first = table + 5ul * njobs + 1ul;\n"
       << "  const unsigned long* pos = first + gridDim.x + 1ul;\n"
       << "  const unsigned long tend = first[blockIdx.x + 1u];\n"
       << "  unsigned long p = pos[blockIdx.x];\n"
       << "  unsigned long b = table[p], nb = table[p + 1ul];\n"
       << "  const unsigned long* rec = table + njobs + 1ul + 4ul * p;\n"
       << "  for (unsigned long tile = first[blockIdx.x]; tile < tend; ++tile) {\n"
       << "    if (tile >= nb) {\n"
       << "      do { ++p; b = nb; nb = table[p + 1ul]; } while (tile >= nb);\n"
       << "      rec = table + njobs + 1ul + 4ul * p;\n"
       << "    }\n"
       << "    const unsigned long lt = tile - b;\n"
       << "    const " << T2 << "* jin = in + rec[0];\n"
       << "    " << T2 << "* jout = out + rec[1];\n"
       << "    const unsigned long K = rec[2];\n"
       // The work-item id is made opaque per tile so that no variant's index arithmetic is
       // hoisted out of the tile loop (with every variant inlined, hoisting would keep all of
       // their invariants live at once).
       << "    unsigned lid = threadIdx.x;\n"
       << "    __asm__ volatile(\"\" : \"+v\"(lid));\n"
       << "    switch ((unsigned)rec[3]) {\n";
    for (std::size_t v = 0; v < rk.variants.size(); ++v) {
        os << "    case " << v << "u: " << variant_name(n, v)
           << "(jin, jout, K, lt, lds_buf, lid); break;\n";
    }
    os << "    default: break;\n"
       << "    }\n"
       << "    __syncthreads();\n"
       << "  }\n"
       << "}\n";
}

} // namespace bbfft
