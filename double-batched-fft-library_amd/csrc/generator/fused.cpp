// Two consecutive passes of a multi-dimensional plan in one kernel (plane kept in LDS).
//
// Replaces, where the plane fits, two of the chained 1-D passes of the reference's nd_fft
// (src/common/algorithm/nd_fft.hpp:52-99), which each make a full round trip to memory.
#include "bbfft/bad_configuration.hpp"
#include "bbfft/detail/generator_impl.hpp"
#include "emitter.hpp"
#include "heuristics.hpp"

#include <algorithm>
#include <cstdlib>
#include <iomanip>
#include <ostream>
#include <sstream>
#include <vector>

namespace bbfft {

namespace {

std::uint64_t fnv1a64(std::string const &s) {
    std::uint64_t h = 1469598103934665603ull;
    for (unsigned char ch : s) {
        h ^= ch;
        h *= 1099511628211ull;
    }
    return h;
}

/// Factorisations of n into at most 3 radices in [2, rmax], largest radix first.
void factorizations(std::size_t n, std::size_t rmax, std::size_t parts, std::vector<int> &cur,
                    std::vector<std::vector<int>> &out) {
    if (n == 1) {
        if (!cur.empty()) {
            out.push_back(cur);
        }
        return;
    }
    if (parts == 0) {
        return;
    }
    std::size_t const hi = cur.empty() ? rmax : std::min<std::size_t>(rmax, cur.back());
    for (std::size_t r = std::min(hi, n); r >= 2; --r) {
        if (n % r == 0) {
            cur.push_back(static_cast<int>(r));
            factorizations(n / r, rmax, parts - 1, cur, out);
            cur.pop_back();
        }
    }
}

/// Kernel configuration of pass `p` restricted to one plane (`kcols` K-columns, one tile):
/// the factorisation with the smallest largest radix (fewest registers per work-item) whose
/// work-group fits `wg_max`; false if none fits.
bool one_tile(configuration p, std::size_t kcols, device_info const &info, unsigned flags,
              unsigned variant, std::size_t wg_max, fft1d_configuration &out) {
    p.shape[2] = kcols;
    std::size_t const M = p.shape[0];
    bool const real = p.type != transform_type::c2c;
    std::size_t const N = p.shape[1];
    std::size_t const nfft = real && N % 2 == 0 ? N / 2 : N;
    std::size_t const cols = real && N % 2 == 1 ? (kcols + 1) / 2 : kcols;
    std::size_t const rmax = p.fp == precision::f32 ? 32 : 16;
    std::vector<std::vector<int>> cands;
    std::vector<int> cur;
    if (nfft <= rmax) {
        cands.push_back({static_cast<int>(nfft)});
    }
    factorizations(nfft, rmax, 3, cur, cands);
    std::vector<int> const *best = nullptr;
    std::size_t best_key = ~std::size_t(0);
    for (auto const &f : cands) {
        std::size_t const r = static_cast<std::size_t>(f.front()); // largest radix
        if (M * (nfft / r) * cols > wg_max) {
            continue;
        }
        // smallest largest radix, then fewest stages, then most balanced
        std::size_t const key = r * 64 + f.size() * 16 + (r - static_cast<std::size_t>(f.back()));
        if (key < best_key) {
            best_key = key;
            best = &f;
        }
    }
    if (!best) {
        return false;
    }
    fft1d_tuning t;
    t.factorization = *best;
    t.Nt = nfft / static_cast<std::size_t>(best->front());
    t.Mb = M;
    t.Kb = cols;
    t.flags = flags;
    t.has_flags = true;
    try {
        out = configure_fft1d(p, info, t, variant);
    } catch (bad_configuration const &) {
        return false;
    }
    return true;
}

/// Complex entries of the twiddle tables of pass c (stages 1.. and the real pre/post-pass),
/// with `pad` the rows of the stage tables one entry apart more.
std::size_t twiddle_entries(fft1d_configuration const &c, bool pad) {
    std::size_t n = 0, p = c.factorization.empty() ? 1 : c.factorization[0];
    for (std::size_t s = 1; s < c.factorization.size(); ++s) {
        n += p * (c.factorization[s] + (pad ? 1 : 0));
        p *= c.factorization[s];
    }
    if (c.type != transform_type::c2c && c.N % 2 == 0) {
        n += c.N / 4 + 1;
    }
    return n;
}

/// Value of environment variable `name` ("" when unset).
std::string env(char const *name) {
    char const *v = std::getenv(name);
    return v ? std::string(v) : std::string();
}

} // namespace

std::string fused2_configuration::identifier() const {
    std::ostringstream oss;
    oss << "bbfft_fused2_f" << static_cast<int>(a.fp) * 8 << "_G" << group_size << "_"
        << std::hex << std::setw(16) << std::setfill('0')
        << fnv1a64(a.identifier() + ";" + b.identifier() + ";" + std::to_string(ka) + ";" +
                   std::to_string(kb) + ";" + std::to_string(plane_elems) + ";" +
                   std::to_string(in_plane_stride) + ";" + std::to_string(out_plane_stride) +
                   ";" + std::to_string(stage_words) + ";" + std::to_string(stage_load_bytes) +
                   ";" + std::to_string(lds_barriers) + ";" + std::to_string(lds_tables) +
                   ";" + std::to_string(waves_per_eu) + ";" + std::to_string(copy_wave0) + ";" +
                   std::to_string(early_copy) + ";" + std::to_string(stage_swizzle_rows) + ";" +
                   std::to_string(stage_swizzle_elems) + ";" + std::to_string(stage_load_nt) + ";" +
                   std::to_string(lds_table_pad));
    return oss.str();
}

std::size_t fused2_configuration::lds_bytes() const {
    std::size_t const cb = 2 * static_cast<std::size_t>(a.fp);
    std::size_t const tables =
        lds_tables ? (twiddle_entries(a, lds_table_pad) + twiddle_entries(b, lds_table_pad)) * cb
                   : 0;
    return std::max({plane_elems * cb, a.lds_bytes(), b.lds_bytes()}) + tables;
}

bool configure_fused2(configuration const &pa, configuration const &pb, device_info const &info,
                      unsigned variant_a, unsigned variant_b, fused2_configuration &out) {
    if (pa.dim != 1 || pb.dim != 1 || pa.fp != pb.fp || pa.callbacks || pb.callbacks) {
        return false;
    }
    if (pa.type == transform_type::c2r || pb.type == transform_type::r2c) {
        return false; // the intermediate must be complex
    }
    std::size_t const Ma = pa.shape[0], Na = pa.shape[1];
    std::size_t const Mb = pb.shape[0], Nb = pb.shape[1];
    std::size_t const na_out = pa.type == transform_type::r2c ? Na / 2 + 1 : Na;
    std::size_t const nb_in = pb.type == transform_type::c2r ? Nb / 2 + 1 : Nb;
    // packed intermediate: pass a writes {1, Ma, Ma na_out}, pass b reads {1, Mb, Mb nb_in}
    if (pa.ostride[0] != 1 || pa.ostride[1] != Ma || pa.ostride[2] != Ma * na_out ||
        pb.istride[0] != 1 || pb.istride[1] != Mb || pb.istride[2] != Mb * nb_in) {
        return false;
    }
    std::size_t const P = std::max(pa.ostride[2], pb.istride[2]);
    if (P % pa.ostride[2] || P % pb.istride[2]) {
        return false;
    }
    std::size_t const ka = P / pa.ostride[2], kb = P / pb.istride[2];
    if (pa.shape[2] % ka || pb.shape[2] % kb || pa.shape[2] / ka != pb.shape[2] / kb) {
        return false;
    }
    fused2_configuration fc = {};
    // As many work-items as the device allows, so each holds few elements (smallest radices);
    // streaming hints on the global side only.
    std::size_t wg_max = std::min<std::size_t>(1024, info.max_work_group_size);
    if (!env("BBFFT_FUSE_WGMAX").empty()) { // A/B: smaller groups (several per CU)
        wg_max = std::min<std::size_t>(
            wg_max, static_cast<std::size_t>(std::atoi(env("BBFFT_FUSE_WGMAX").c_str())));
    }
    if (!one_tile(pa, ka, info, fft1d_nontemporal_load, variant_a, wg_max, fc.a) ||
        !one_tile(pb, kb, info, fft1d_nontemporal_store, variant_b, wg_max, fc.b)) {
        return false;
    }
    fc.ka = ka;
    fc.kb = kb;
    fc.planes = pa.shape[2] / ka;
    fc.plane_elems = P;
    fc.in_plane_stride = ka * pa.istride[2];
    fc.out_plane_stride = kb * pb.ostride[2];
    fc.group_size = std::max(fc.a.work_group_size(), fc.b.work_group_size());
    // Kernel variants (all compute the same bits). Measured on MI355X with bench.py's 3-D legs
    // (32 x 128^3 fp32, whole transform, processes alternated, 3 rounds): r2c 0.215 ms with
    // barriers ordering LDS only, LDS twiddle tables, 16-byte staging loads issued right after
    // pass a's first stage by the waves idle in pass b, against 0.225 without the idle-wave
    // split, 0.234 without the early copy, 0.234 with 4-byte loads and 0.232 with none of
    // these; c2r (c2c first pass, all work-items busy in pass b) 0.229 ms with 16-byte loads
    // alone, against 0.233 with 4-byte loads, 0.233-0.234 with LDS-only barriers and
    // (unpadded) tables and 0.238 with the early copy as well; with padded table rows the
    // LDS-only barriers and tables became the faster c2r setting (below). Environment
    // overrides for A/B: BBFFT_FUSE_SYNC=full|lds, BBFFT_FUSE_TABLES=const|lds,
    // BBFFT_FUSE_TABPAD=on|off, BBFFT_FUSE_DMA=4|16, BBFFT_FUSE_EARLY=off|on,
    // BBFFT_FUSE_DMAWAVES=all|split|idle_a, BBFFT_FUSE_DMA_NT=on|off,
    // BBFFT_FUSE_WAVES=<waves per SIMD>.
    bool const r2c_first = pa.type == transform_type::r2c;
    // LDS-only barriers and twiddle tables in LDS (rows padded, below) for real transforms;
    // with the padded rows the c2r planes gained too (32 x 128^3 c2r 0.2027 ms against 0.2061,
    // 2048 x 128^2 c2r 0.0557 against 0.0579), while c2c planes lost (32 x 64 x 128^2
    // 0.2613 / 0.2618 ms against 0.2567 / 0.2573 forward / backward).
    bool const real_plan = r2c_first || pb.type == transform_type::c2r;
    auto option = [](char const *name, char const *on, bool dflt) {
        std::string const v = env(name);
        return v.empty() ? dflt : v == on;
    };
    fc.lds_barriers = option("BBFFT_FUSE_SYNC", "lds", real_plan);
    fc.lds_tables = option("BBFFT_FUSE_TABLES", "lds", real_plan);
    // LDS twiddle rows padded to R + 1 entries (BBFFT_FUSE_TABPAD=off for A/B): the first
    // pass's work-items t and t + 4 read rows 64 bytes apart, on the same banks. 32 x 128^3
    // r2c 0.1974 ms against 0.1993 unpadded (processes alternated).
    fc.lds_table_pad = fc.lds_tables && option("BBFFT_FUSE_TABPAD", "on", true);
    if (!env("BBFFT_FUSE_WAVES").empty()) {
        fc.waves_per_eu = static_cast<unsigned>(std::atoi(env("BBFFT_FUSE_WAVES").c_str()));
    }
    if (fc.lds_bytes() > info.local_memory_size) {
        fc.lds_tables = false;
        fc.lds_table_pad = false;
    }
    if (fc.lds_bytes() > info.local_memory_size) {
        return false;
    }
    // Staging: with the plane's raw input in a second LDS buffer, the next plane's input loads
    // (direct global-to-LDS, 4 bytes per work-item and instruction) run while the current
    // plane's second pass computes and stores. Needs room for both buffers; BBFFT_FUSE_STAGE=off
    // disables it.
    char const *stage_env = std::getenv("BBFFT_FUSE_STAGE");
    bool const stage_off = stage_env && std::string(stage_env) == "off";
    std::size_t const in_bytes = pa.type == transform_type::r2c
                                     ? static_cast<std::size_t>(pa.fp)
                                     : 2 * static_cast<std::size_t>(pa.fp);
    std::size_t const words = (fc.in_plane_stride * in_bytes / 4 + 63) / 64 * 64;
    // the plane's input must lie within its in_plane_stride elements (it does for the
    // default packed and in-place layouts that multi-dimensional plans use)
    std::size_t const na_in = pa.shape[1];
    std::size_t const extent = (Ma - 1) * pa.istride[0] + (na_in - 1) * pa.istride[1] +
                               (ka - 1) * pa.istride[2] + 1;
    if (!stage_off && info.type == device_type::gpu && extent <= fc.in_plane_stride &&
        fc.lds_bytes() + 4 * words <= info.local_memory_size) {
        fused2_configuration st = fc;
        // pass a now reads LDS: no streaming hint on its loads
        if (one_tile(pa, ka, info, 0u, variant_a, wg_max, st.a) &&
            st.a.work_group_size() == fc.a.work_group_size() &&
            st.lds_bytes() + 4 * words <= info.local_memory_size) {
            st.stage_words = words;
            // whole 1 KB blocks per wave when the plane's input is a multiple of 16 bytes
            // (BBFFT_FUSE_DMA=4 keeps one word per work-item)
            std::size_t const W = fc.in_plane_stride * in_bytes / 4;
            st.stage_load_bytes = W % 4 == 0 && fc.group_size >= 64 &&
                                          option("BBFFT_FUSE_DMA", "16", true)
                                      ? 16
                                      : 4;
            // Staging loads with the non-temporal cache policy (BBFFT_FUSE_DMA_NT=on); plans
            // choose it from their other passes' stores (find_fused_pair).
            st.stage_load_nt = option("BBFFT_FUSE_DMA_NT", "on", false);
            // staging loads from the waves idle in pass b (BBFFT_FUSE_DMAWAVES=split) or, with
            // the early copy, from those idle in pass a after its first stage (=idle_a)
            std::size_t const a_waves = (st.a.work_group_size() + 63) / 64;
            std::size_t const b_waves = (st.b.work_group_size() + 63) / 64;
            std::string const dma_waves = env("BBFFT_FUSE_DMAWAVES");
            bool const lds_only = st.lds_barriers && st.lds_tables && st.group_size % 64 == 0;
            if ((dma_waves.empty() ? r2c_first : dma_waves == "split") && lds_only &&
                b_waves < st.group_size / 64) {
                st.copy_wave0 = static_cast<unsigned>(b_waves);
            } else if (dma_waves == "idle_a" && lds_only && a_waves < st.group_size / 64) {
                st.copy_wave0 = static_cast<unsigned>(a_waves);
            }
            // next plane's copy issued right after pass a's first stage (BBFFT_FUSE_EARLY=on);
            // only useful when barriers do not wait for it
            st.early_copy = option("BBFFT_FUSE_EARLY", "on", r2c_first) &&
                            (st.a.factorization.size() > 1 || st.a.type == transform_type::r2c);
            // Rows of pass a read from the stage: a wave holds 64 / Nt rows and reads one
            // segment of Nt complex elements of each per instruction. Rows of a multiple of
            // (64 / Nt) segments alias in the LDS banks; permuting the segments row by row
            // spreads them. Measured (bench.py's 3-D legs, processes alternated): 32 x 128^3
            // r2c 0.2107 ms against 0.2166 with the plain layout; a c2c first pass
            // (32 x 128 x 128 x 64) 0.2266 against 0.2220, so r2c first passes only
            // (BBFFT_FUSE_SWIZZLE=on|off overrides).
            std::size_t const Nt = st.a.Nt;
            std::size_t const eb = 2 * static_cast<std::size_t>(pa.fp); // complex element
            bool const complex_rows =
                pa.istride[1] == 1 && !pa.callbacks &&
                (pa.type == transform_type::c2c ||
                 (pa.type == transform_type::r2c && Na % 2 == 0 && pa.istride[2] % 2 == 0));
            std::size_t const row_bytes = pa.istride[2] * in_bytes;
            if (option("BBFFT_FUSE_SWIZZLE", "on", r2c_first) && complex_rows && Ma == 1 &&
                st.a.Mb == 1 && Nt >= 2 && Nt <= 32 && 64 % Nt == 0 && (Nt * eb) % 16 == 0 &&
                row_bytes % ((64 / Nt) * Nt * eb) == 0 && W % 4 == 0) {
                st.stage_swizzle_rows = static_cast<unsigned>(64 / Nt);
                st.stage_swizzle_elems = static_cast<unsigned>(Nt);
            }
            fc = st;
        }
    }
    out = fc;
    return true;
}

void generate_fused2(std::ostream &os, fused2_configuration const &fc, std::string_view name) {
    std::string const n = name.empty() ? fc.identifier() : std::string(name);
    bool const f32 = fc.a.fp == precision::f32;
    std::string const T = f32 ? "float" : "double";
    std::string const T2 = f32 ? "float2" : "double2";
    std::string const in_t = fc.a.type == transform_type::r2c ? T : T2;
    std::string const out_t = fc.b.type == transform_type::c2r ? T : T2;
    std::size_t const lds_elems = fc.lds_bytes() / (2 * static_cast<std::size_t>(fc.a.fp));
    // body names derive from the kernel name: several kernels may share one source
    std::string const fa = n + "_a", fb = n + "_b";
    if (fc.lds_barriers) {
        // Barrier ordering LDS only: waits for this wave's outstanding LDS operations, not for
        // global loads and stores, nor for the direct global-to-LDS staging loads (a fence on
        // LDS would wait for those, as they write LDS; the staging buffer is only read after
        // BBFFT_WAIT_ALL). The memory clobber keeps the compiler from moving LDS accesses
        // across it.
        // (Every kernel source that defines it defines it the same way, so kernels compiled
        // together in one translation unit agree; kernels without lds_barriers use
        // __syncthreads().)
        os << "#ifndef BBFFT_LDS_SYNC\n"
              "#define BBFFT_LDS_SYNC() asm volatile(\"s_waitcnt lgkmcnt(0)\\n\\ts_barrier\" ::: "
              "\"memory\")\n"
              "#endif\n";
    }
    os << hip_prelude();
    if (fc.stage_words) {
        // direct global-to-LDS load of one 32-bit word per work-item (the LDS address is the
        // wave-uniform base plus 4 bytes per lane) and a wait for all outstanding loads
        os << "#ifndef BBFFT_LDS_DMA4\n"
              "#define BBFFT_LDS_DMA4(src, dst) __builtin_amdgcn_global_load_lds("
              "(const __attribute__((address_space(1))) void*)(src), "
              "(__attribute__((address_space(3))) void*)(dst), 4, 0, 0)\n"
              "#define BBFFT_WAIT_ALL() __builtin_amdgcn_s_waitcnt(0)\n"
              "#endif\n";
        if (fc.stage_load_bytes == 16) {
            os << "#ifndef BBFFT_LDS_DMA16\n"
                  "#define BBFFT_LDS_DMA16(src, dst) __builtin_amdgcn_global_load_lds("
                  "(const __attribute__((address_space(1))) void*)(src), "
                  "(__attribute__((address_space(3))) void*)(dst), 16, 0, 0)\n"
                  "#endif\n";
        }
        if (fc.stage_load_nt) {
            // the same loads with the non-temporal cache policy (aux bit 1 on gfx950)
            os << "#ifndef BBFFT_LDS_DMA4_NT\n"
                  "#define BBFFT_LDS_DMA4_NT(src, dst) __builtin_amdgcn_global_load_lds("
                  "(const __attribute__((address_space(1))) void*)(src), "
                  "(__attribute__((address_space(3))) void*)(dst), 4, 0, 2)\n"
                  "#define BBFFT_LDS_DMA16_NT(src, dst) __builtin_amdgcn_global_load_lds("
                  "(const __attribute__((address_space(1))) void*)(src), "
                  "(__attribute__((address_space(3))) void*)(dst), 16, 0, 2)\n"
                  "#endif\n";
        }
    }
    std::size_t const in_bytes = fc.a.type == transform_type::r2c
                                     ? static_cast<std::size_t>(fc.a.fp)
                                     : 2 * static_cast<std::size_t>(fc.a.fp);
    std::size_t const W = fc.in_plane_stride * in_bytes / 4; // words of one plane's input
    std::string const stage = n + "_stage", copy = n + "_copy";
    if (fc.stage_words) {
        // Staged: plane f's raw input (W words) sits in `stage` before pass a starts; the copy
        // of the work-group's next plane is issued once pass a no longer reads `stage` (after
        // pass a, or with early_copy right after pass a's first stage) and overlaps the rest.
        // Blocks of BS words, one per copying wave and step: lane l's word lands at the block
        // base + l; a group smaller than one wave copies group_size words per step.
        std::size_t const BS = std::min<std::size_t>(64, fc.group_size);
        std::size_t const wv0 = fc.copy_wave0; // copying waves: wv0 .. wv0 + NWc - 1
        std::size_t const NWc = fc.group_size >= 64 ? fc.group_size / 64 - wv0 : 1;
        std::size_t const NB = (W + BS - 1) / BS;
        // Source word of stage word i: i itself, or with stage_swizzle_rows the word of the
        // same row whose 16-byte chunk index is XOR-ed with (row % R) * (segment chunks).
        std::string src_word = "i";
        if (fc.stage_swizzle_rows) {
            std::size_t const in_b = fc.a.type == transform_type::r2c
                                         ? static_cast<std::size_t>(fc.a.fp)
                                         : 2 * static_cast<std::size_t>(fc.a.fp);
            std::size_t const CR = fc.a.istride[2] * in_b / 16; // 16-byte chunks per row
            std::size_t const SC = fc.stage_swizzle_elems * 2 * static_cast<std::size_t>(fc.a.fp) / 16;
            src_word = "((i / " + std::to_string(4 * CR) + "u) * " + std::to_string(4 * CR) +
                       "u + ((((i / 4u) % " + std::to_string(CR) + "u) ^ (((i / " +
                       std::to_string(4 * CR) + "u) % " + std::to_string(fc.stage_swizzle_rows) +
                       "u) * " + std::to_string(SC) + "u)) * 4u) + i % 4u)";
        }
        std::string const nt = fc.stage_load_nt ? "_NT" : "";
        std::string const dma16 = "BBFFT_LDS_DMA16" + nt, dma4 = "BBFFT_LDS_DMA4" + nt;
        os << "// copy of one plane's input (src, or nothing when src is null) into the staging\n"
           << "// buffer `stage` (LDS)\n"
           << "static __device__ __forceinline__ void " << copy
           << "(const float* src, float* stage) {\n"
           << "  const unsigned ln = threadIdx.x % 64u;\n"
           << "  const unsigned wv = threadIdx.x / 64u - " << wv0 << "u; // copying wave index\n"
           << "  if (!src || wv >= " << NWc << "u) return; // other waves copy nothing\n";
        if (fc.stage_swizzle_rows) {
            os << "  // rows permuted in 16-byte chunks: " << fc.stage_swizzle_rows
               << " rows read together land in different LDS banks\n";
        }
        if (fc.stage_load_bytes == 16) {
            // 16 bytes per work-item: wave block b of 256 words lands at stage + 256 b, lane
            // l's four words at + 4 l (the pointer's alignment is uniform: one branch)
            std::size_t const NB16 = (W + 255) / 256;
            os << "  if ((reinterpret_cast<unsigned long>(src) & 15ul) == 0ul) {\n"
               << "    for (unsigned b = wv; b < " << NB16 << "u; b += " << NWc << "u) {\n"
               << "      const unsigned i = b * 256u + 4u * ln;\n"
               << "      if (i < " << W << "u) " << dma16 << "(src + " << src_word
               << ", stage + b * 256u);\n"
               << "    }\n"
               << "    return;\n"
               << "  }\n";
        }
        os << "  for (unsigned b = wv; b < " << NB << "u; b += " << NWc << "u) {\n"
           << "    const unsigned i = b * " << BS << "u + ln;\n"
           << "    if (i < " << W << "u) " << dma4 << "(src + " << src_word
           << ", stage + b * " << BS << "u);\n"
           << "  }\n"
           << "}\n";
    }
    body_options oa, ob;
    oa.lds_tables = ob.lds_tables = fc.lds_tables;
    oa.lds_table_pad = ob.lds_table_pad = fc.lds_table_pad;
    oa.lds_barriers = ob.lds_barriers = fc.lds_barriers;
    oa.single_tile = ob.single_tile = true; // one tile per plane: Kb = K, Mb = M
    std::string const sync = fc.lds_barriers ? "BBFFT_LDS_SYNC();" : "__syncthreads();";
    if (fc.stage_words && fc.early_copy) {
        oa.extra_param = "const float* next_input, float* stage";
        oa.after_input = copy + "(next_input, stage);";
    }
    if (fc.stage_words && fc.stage_swizzle_rows) {
        oa.input_index_xor = "(" + std::to_string(fc.stage_swizzle_elems) + "u * (kl % " +
                             std::to_string(fc.stage_swizzle_rows) + "u))";
    }
    generate_fft1d_body_aliased(os, fc.a, fa, fc.group_size, false, true, oa);
    generate_fft1d_body_aliased(os, fc.b, fb, fc.group_size, true, false, ob);
    os << "extern \"C\" __global__ void __launch_bounds__(" << fc.group_size << ") ";
    if (fc.waves_per_eu) {
        os << "__attribute__((amdgpu_waves_per_eu(" << fc.waves_per_eu << "))) ";
    }
    os << n
       << "(const " << in_t << "* in, " << out_t << "* out, unsigned long planes) {\n"
       << "  __shared__ " << T2 << " lds_buf[" << lds_elems << "];\n";
    if (fc.lds_tables) {
        // the first barrier of the plane loop orders these writes before any table read
        os << "  " << fa << "_tables(threadIdx.x, " << fc.group_size << "u);\n"
           << "  " << fb << "_tables(threadIdx.x, " << fc.group_size << "u);\n";
    }
    if (!fc.stage_words) {
        os << "  for (unsigned long f = blockIdx.x; f < planes; f += gridDim.x) {\n"
           << "    " << sync << "\n"
           << "    " << fa << "(in + f * " << fc.in_plane_stride << "ul, lds_buf, " << fc.ka
           << "ul, 0ul, lds_buf, threadIdx.x);\n"
           << "    " << sync << "\n"
           << "    " << fb << "(lds_buf, out + f * " << fc.out_plane_stride << "ul, " << fc.kb
           << "ul, 0ul, lds_buf, threadIdx.x);\n"
           << "  }\n"
           << "}\n";
        return;
    }
    std::size_t const NWc = fc.group_size >= 64 ? fc.group_size / 64 - fc.copy_wave0 : 1;
    std::string const next = "(f + gridDim.x < planes ? (const float*)in + (f + gridDim.x) * " +
                             std::to_string(W) + "ul : nullptr)";
    // 16-byte aligned: pass a may read the staged input as float2 / double2
    os << "  __shared__ __attribute__((aligned(16))) float " << stage << "[" << fc.stage_words
       << "];\n"
       << "  unsigned long f = blockIdx.x;\n"
       << "  if (f < planes) " << copy << "((const float*)in + f * " << W << "ul, " << stage
       << ");\n"
       << "  for (; f < planes; f += gridDim.x) {\n";
    if (fc.copy_wave0) {
        os << "    if (threadIdx.x / 64u - " << fc.copy_wave0 << "u < " << NWc
           << "u) BBFFT_WAIT_ALL(); // copying waves\n";
    } else {
        os << "    BBFFT_WAIT_ALL();\n";
    }
    os << "    " << sync << "\n"
       << "    " << fa << "((const " << in_t << "*)" << stage << ", lds_buf, " << fc.ka
       << "ul, 0ul, lds_buf, threadIdx.x"
       << (fc.early_copy ? ", " + next + ", " + stage : std::string()) << ");\n"
       << "    " << sync << "\n";
    if (!fc.early_copy) {
        os << "    " << copy << "(" << next << ", " << stage << ");\n";
    }
    os << "    " << fb << "(lds_buf, out + f * " << fc.out_plane_stride << "ul, " << fc.kb
       << "ul, 0ul, lds_buf, threadIdx.x);\n"
       << "    " << sync << "\n"
       << "  }\n"
       << "}\n";
}

} // namespace bbfft
