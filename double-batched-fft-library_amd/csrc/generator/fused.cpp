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

/// Complex entries of the twiddle tables of pass c (stages 1.. and the real pre/post-pass).
std::size_t twiddle_entries(fft1d_configuration const &c) {
    std::size_t n = 0, p = c.factorization.empty() ? 1 : c.factorization[0];
    for (std::size_t s = 1; s < c.factorization.size(); ++s) {
        n += p * c.factorization[s];
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
                   ";" + std::to_string(waves_per_eu));
    return oss.str();
}

std::size_t fused2_configuration::lds_bytes() const {
    std::size_t const cb = 2 * static_cast<std::size_t>(a.fp);
    std::size_t const tables = lds_tables ? (twiddle_entries(a) + twiddle_entries(b)) * cb : 0;
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
    std::size_t const wg_max = std::min<std::size_t>(1024, info.max_work_group_size);
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
    // Barriers and tables (BBFFT_FUSE_SYNC=full|lds, BBFFT_FUSE_TABLES=const|lds)
    fc.lds_barriers = env("BBFFT_FUSE_SYNC") == "lds";
    fc.lds_tables = env("BBFFT_FUSE_TABLES") == "lds";
    if (!env("BBFFT_FUSE_WAVES").empty()) {
        fc.waves_per_eu = static_cast<unsigned>(std::atoi(env("BBFFT_FUSE_WAVES").c_str()));
    }
    if (fc.lds_bytes() > info.local_memory_size) {
        fc.lds_tables = false;
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
            st.stage_load_bytes =
                W % 4 == 0 && fc.group_size >= 64 && env("BBFFT_FUSE_DMA") == "16" ? 16 : 4;
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
        os << "#ifndef BBFFT_SYNC\n"
              "#define BBFFT_SYNC() asm volatile(\"s_waitcnt lgkmcnt(0)\\n\\ts_barrier\" ::: "
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
    }
    generate_fft1d_body_aliased(os, fc.a, fa, fc.group_size, false, true, fc.lds_tables);
    generate_fft1d_body_aliased(os, fc.b, fb, fc.group_size, true, false, fc.lds_tables);
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
           << "    BBFFT_SYNC();\n"
           << "    " << fa << "(in + f * " << fc.in_plane_stride << "ul, lds_buf, " << fc.ka
           << "ul, 0ul, lds_buf, threadIdx.x);\n"
           << "    BBFFT_SYNC();\n"
           << "    " << fb << "(lds_buf, out + f * " << fc.out_plane_stride << "ul, " << fc.kb
           << "ul, 0ul, lds_buf, threadIdx.x);\n"
           << "  }\n"
           << "}\n";
        return;
    }
    // Staged: plane f's raw input (W words) sits in `stage` before pass a starts; the copy of
    // plane f + gridDim.x is issued once pass a has consumed `stage` and overlaps pass b.
    // Each wave copies blocks of 64 words: lane l's word lands at the block base + l.
    std::size_t const in_bytes = fc.a.type == transform_type::r2c
                                     ? static_cast<std::size_t>(fc.a.fp)
                                     : 2 * static_cast<std::size_t>(fc.a.fp);
    std::size_t const W = fc.in_plane_stride * in_bytes / 4;
    // Blocks of BS words, one per copying wave and step: the full waves copy 64 words each;
    // a group smaller than one wave copies group_size words per step with its only wave.
    std::size_t const BS = std::min<std::size_t>(64, fc.group_size);
    std::size_t const NWc = fc.group_size >= 64 ? fc.group_size / 64 : 1;
    std::size_t const NB = (W + BS - 1) / BS;
    // 16-byte aligned: pass a may read the staged input as float2 / double2
    os << "  __shared__ __attribute__((aligned(16))) float stage[" << fc.stage_words << "];\n"
       << "  const unsigned wv = threadIdx.x / 64u, ln = threadIdx.x % 64u;\n"
       << "  auto stage_copy = [&](unsigned long f) {\n"
       << "    const float* src = (const float*)in + f * " << W << "ul;\n"
       << "    if (wv >= " << NWc << "u) return; // a partial last wave copies nothing\n";
    if (fc.stage_load_bytes == 16) {
        // 16 bytes per work-item: wave block b of 256 words lands at stage + 256 b, lane l's
        // four words at + 4 l (the input pointer's alignment is uniform: one branch)
        std::size_t const NB16 = (W + 255) / 256;
        os << "    if ((reinterpret_cast<unsigned long>(in) & 15ul) == 0ul) {\n"
           << "      for (unsigned b = wv; b < " << NB16 << "u; b += " << NWc << "u) {\n"
           << "        const unsigned i = b * 256u + 4u * ln;\n"
           << "        if (i < " << W << "u) BBFFT_LDS_DMA16(src + i, stage + b * 256u);\n"
           << "      }\n"
           << "      return;\n"
           << "    }\n";
    }
    os << "    for (unsigned b = wv; b < " << NB << "u; b += " << NWc << "u) {\n"
       << "      const unsigned i = b * " << BS << "u + ln;\n"
       << "      if (i < " << W << "u) BBFFT_LDS_DMA4(src + i, stage + b * " << BS << "u);\n"
       << "    }\n"
       << "  };\n"
       << "  unsigned long f = blockIdx.x;\n"
       << "  if (f < planes) stage_copy(f);\n"
       << "  for (; f < planes; f += gridDim.x) {\n"
       << "    BBFFT_WAIT_ALL();\n"
       << "    BBFFT_SYNC();\n"
       << "    " << fa << "((const " << in_t << "*)stage, lds_buf, " << fc.ka
       << "ul, 0ul, lds_buf, threadIdx.x);\n"
       << "    BBFFT_SYNC();\n"
       << "    if (f + gridDim.x < planes) stage_copy(f + gridDim.x);\n"
       << "    " << fb << "(lds_buf, out + f * " << fc.out_plane_stride << "ul, " << fc.kb
       << "ul, 0ul, lds_buf, threadIdx.x);\n"
       << "  }\n"
       << "}\n";
}

} // namespace bbfft
