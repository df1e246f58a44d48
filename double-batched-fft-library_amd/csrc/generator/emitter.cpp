// HIP source generator for gfx950 double-batched 1-D FFT kernels.
//
// Replaces the reference's clir/OpenCL-C generators (src/base/generator/sbfft_gen.cpp:30-351,
// f2fft_gen.cpp:46-502, mixed_radix_fft.cpp:25-266). The kernel structure is designed for CDNA:
//   * work-item (ml, t, kl) of a Mb x Nt x Kb work-group; the Mb M-columns are the fastest index
//     so a wave's global accesses are contiguous along the stride-1 M mode;
//   * Stockham autosort stages: stage s with radix R reads x[j + i N/R] and writes
//     y[(j / p) p R + j % p + i p] (p = product of earlier radices); the first stage reads global
//     memory and the last stage writes global memory, both coalesced across j; intermediate
//     stages go through padded LDS;
//   * radix-R butterflies are straight-line register code: Cooley-Tukey splits with radix-4/2
//     kernels, symmetric-pair odd-prime kernels and constant twiddles simplified at generation;
//   * real transforms of even length use the half-length complex FFT with a fused LDS pre/post
//     pass (reference: sbfft_gen.cpp:180-247); odd lengths pack two K-columns into one complex FFT
//     (reference: sbfft_gen.cpp:249-351);
//   * callbacks are plain HIP device functions called with element offsets
//     (reference: src/base/generator/tensor_accessor.cpp:39-50).
#include "../base/math.hpp"
#include "bbfft/bad_configuration.hpp"
#include "bbfft/detail/generator_impl.hpp"
#include "emitter.hpp"
#include "roots.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ostream>
#include <sstream>
#include <string>
#include <vector>

namespace bbfft {

namespace {

/// Round to the nearest value with `mant` fraction bits and minimum normal exponent `emin`
/// (ties to even); subnormals keep the quantum 2^(emin - mant). No overflow handling: the
/// values rounded here have magnitude <= 1.
double round_to_format(double x, int mant, int emin) {
    if (x == 0.0) {
        return x;
    }
    int e = 0;
    std::frexp(x, &e); // |x| = f 2^e, f in [0.5, 1)
    int const ex = std::max(e - 1, emin);
    double const quantum = std::ldexp(1.0, ex - mant);
    return std::nearbyint(x / quantum) * quantum; // default rounding mode: ties to even
}
double round_f16(double x) { return round_to_format(x, 10, -14); }
double round_bf16(double x) { return round_to_format(x, 7, -126); }

/// IEEE binary16 bit pattern of a value representable in binary16.
std::uint16_t f16_bits(double v) {
    std::uint16_t sign = std::signbit(v) ? 0x8000u : 0u;
    double a = std::fabs(v);
    if (a == 0.0) {
        return sign;
    }
    if (a < std::ldexp(1.0, -14)) {
        return static_cast<std::uint16_t>(sign | static_cast<std::uint16_t>(a * std::ldexp(1.0, 24)));
    }
    int e = 0;
    double f = std::frexp(a, &e); // a = f 2^e
    unsigned const exp = static_cast<unsigned>(e - 1 + 15);
    unsigned const frac = static_cast<unsigned>(f * 2048.0) - 1024u;
    return static_cast<std::uint16_t>(sign | (exp << 10) | frac);
}
/// bfloat16 bit pattern of a value representable in bfloat16.
std::uint16_t bf16_bits(double v) {
    float f = static_cast<float>(v); // exact
    std::uint32_t b = 0;
    std::memcpy(&b, &f, sizeof(b));
    return static_cast<std::uint16_t>(b >> 16);
}

/// Types and conversions for kernels with half-precision input (emitted when needed).
std::string half_prelude() {
    return "#ifndef BBFFT_HIP_HALF_PRELUDE\n#define BBFFT_HIP_HALF_PRELUDE\n"
           "typedef _Float16 bbfft_h8 __attribute__((ext_vector_type(8)));\n"
           "typedef __bf16 bbfft_b8 __attribute__((ext_vector_type(8)));\n"
           "typedef unsigned bbfft_u4 __attribute__((ext_vector_type(4)));\n"
           "typedef float bbfft_f16v __attribute__((ext_vector_type(16)));\n"
           "static __device__ __forceinline__ float2 f16c_to_f2(unsigned v) { return "
           "mk((float)__builtin_bit_cast(_Float16, (unsigned short)(v & 0xffffu)), "
           "(float)__builtin_bit_cast(_Float16, (unsigned short)(v >> 16))); }\n"
           "static __device__ __forceinline__ float2 bf16c_to_f2(unsigned v) { return "
           "mk(__builtin_bit_cast(float, v << 16), __builtin_bit_cast(float, v & 0xffff0000u)); "
           "}\n"
           "static __device__ __forceinline__ unsigned ld_nt(const unsigned* p) { return "
           "__builtin_nontemporal_load(p); }\n"
           "#endif\n";
}

/// Line-oriented code writer with fresh-name generation.
class writer {
  public:
    void line(std::string const &s) { os_ << std::string(2 * indent_, ' ') << s << '\n'; }
    void open(std::string const &s) {
        line(s + " {");
        ++indent_;
    }
    void close(std::string const &tail = "}") {
        --indent_;
        line(tail);
    }
    std::string fresh(std::string const &prefix) { return prefix + std::to_string(counter_++); }
    std::string str() const { return os_.str(); }

  private:
    std::ostringstream os_;
    int indent_ = 0;
    int counter_ = 0;
};

/// Types and literals of one precision.
struct fp_helper {
    precision fp;
    char const *real() const { return fp == precision::f32 ? "float" : "double"; }
    char const *cplx() const { return fp == precision::f32 ? "float2" : "double2"; }
    std::string lit(double v) const {
        char buf[64];
        if (fp == precision::f32) {
            std::snprintf(buf, sizeof(buf), "%af", static_cast<double>(static_cast<float>(v)));
        } else {
            std::snprintf(buf, sizeof(buf), "%a", v);
        }
        return buf;
    }
    /// Literal pair (re, im) of exp(2 pi i k / n) rounded once from quad precision to T.
    std::pair<std::string, std::string> root(long k, long n) const {
        if (fp == precision::f32) {
            auto w = gen::root_of_unity_f(k, n);
            return {lit(w.real()), lit(w.imag())};
        }
        auto w = gen::root_of_unity(k, n);
        return {lit(w.real()), lit(w.imag())};
    }
    /// Literal of cos / sin (2 pi k / n).
    std::string cos_lit(long k, long n) const { return root(k, n).first; }
    std::string sin_lit(long k, long n) const { return root(k, n).second; }
    std::string zero2() const { return std::string("mk(") + lit(0.0) + ", " + lit(0.0) + ")"; }
};

long modn(long a, long n) {
    long r = a % n;
    return r < 0 ? r + n : r;
}

/// Straight-line DFT code generator working on named complex SSA values.
class dft_gen {
  public:
    dft_gen(writer &w, fp_helper f, int dir) : w_(w), f_(f), dir_(dir) {}

    std::string let(std::string const &expr) {
        auto n = w_.fresh("r");
        w_.line(std::string("const ") + f_.cplx() + " " + n + " = " + expr + ";");
        return n;
    }
    std::string add(std::string const &a, std::string const &b) {
        return let("c_add(" + a + ", " + b + ")");
    }
    std::string sub(std::string const &a, std::string const &b) {
        return let("c_sub(" + a + ", " + b + ")");
    }

    /// x * exp(2 pi i k / n), with trivial and diagonal constants simplified.
    std::string mul_root(std::string const &x, long k, long n) {
        long r = modn(k, n);
        if (r == 0) {
            return x;
        }
        if (2 * r == n) {
            return let("c_neg(" + x + ")");
        }
        if (4 * r == n) {
            return let("c_ipos(" + x + ")");
        }
        if (4 * r == 3 * n) {
            return let("c_ineg(" + x + ")");
        }
        auto [c, s] = f_.root(r, n);
        if ((8 * r) % n == 0) {
            // odd multiple of pi/4: w = s0 (a + i b), a, b = +-1
            long oct = 8 * r / n; // 1, 3, 5, 7
            int a = (oct == 1 || oct == 7) ? 1 : -1;
            int b = (oct == 1 || oct == 3) ? 1 : -1;
            std::string h = f_.lit(0.70710678118654752440);
            // (x.x a - x.y b, x.x b + x.y a) * s0
            std::string re = std::string(a > 0 ? "" : "-") + x + ".x" + (b > 0 ? " - " : " + ") +
                             x + ".y";
            std::string im = std::string(b > 0 ? "" : "-") + x + ".x" + (a > 0 ? " + " : " - ") +
                             x + ".y";
            return let("c_scale(mk(" + re + ", " + im + "), " + h + ")");
        }
        return let("c_mulri(" + x + ", " + c + ", " + s + ")");
    }

    /// y[k] = sum_i x[i] exp(2 pi i dir i k / R), natural order.
    std::vector<std::string> dft(std::vector<std::string> const &x) {
        int const R = static_cast<int>(x.size());
        if (R == 1) {
            return x;
        }
        if (R == 2) {
            return {add(x[0], x[1]), sub(x[0], x[1])};
        }
        if (R == 4) {
            return radix4(x);
        }
        if (is_prime(static_cast<unsigned>(R))) {
            return prime(x);
        }
        int A = R % 4 == 0 ? 4 : trial_division(R).front();
        if (A == R) {
            A = trial_division(R).front();
        }
        return split(x, A, R / A);
    }

  private:
    std::vector<std::string> radix4(std::vector<std::string> const &x) {
        auto t0 = add(x[0], x[2]);
        auto t1 = sub(x[0], x[2]);
        auto t2 = add(x[1], x[3]);
        auto d = sub(x[1], x[3]);
        // w4^dir * d: forward multiplies by -i, backward by +i
        auto t3 = let(std::string(dir_ < 0 ? "c_ineg(" : "c_ipos(") + d + ")");
        return {add(t0, t2), add(t1, t3), sub(t0, t2), sub(t1, t3)};
    }

    /// Odd prime R with the symmetric-pair formula:
    /// y_k = x_0 + sum_j a_j cos(2 pi jk/R) + dir i sum_j b_j sin(2 pi jk/R), a_j = x_j + x_{R-j},
    /// b_j = x_j - x_{R-j}; y_{R-k} flips the sign of the sine sum.
    std::vector<std::string> prime(std::vector<std::string> const &x) {
        int const R = static_cast<int>(x.size());
        int const h = (R - 1) / 2;
        std::vector<std::string> a(h + 1), b(h + 1);
        for (int j = 1; j <= h; ++j) {
            a[j] = add(x[j], x[R - j]);
            b[j] = sub(x[j], x[R - j]);
        }
        std::vector<std::string> y(R);
        {
            std::string re = x[0] + ".x", im = x[0] + ".y";
            for (int j = 1; j <= h; ++j) {
                re += " + " + a[j] + ".x";
                im += " + " + a[j] + ".y";
            }
            y[0] = let("mk(" + re + ", " + im + ")");
        }
        for (int k = 1; k <= h; ++k) {
            std::string cr = x[0] + ".x", ci = x[0] + ".y", sr, si;
            for (int j = 1; j <= h; ++j) {
                long jk = static_cast<long>(j) * k;
                std::string c = f_.cos_lit(jk, R);
                std::string s = f_.sin_lit(jk, R);
                cr += " + " + a[j] + ".x * " + c;
                ci += " + " + a[j] + ".y * " + c;
                sr += std::string(j > 1 ? " + " : "") + b[j] + ".x * " + s;
                si += std::string(j > 1 ? " + " : "") + b[j] + ".y * " + s;
            }
            auto C = let("mk(" + cr + ", " + ci + ")");
            auto S = let("mk(" + sr + ", " + si + ")");
            // y_k = C + dir i S, y_{R-k} = C - dir i S; i S = (-S.y, S.x)
            if (dir_ < 0) {
                y[k] = let("mk(" + C + ".x + " + S + ".y, " + C + ".y - " + S + ".x)");
                y[R - k] = let("mk(" + C + ".x - " + S + ".y, " + C + ".y + " + S + ".x)");
            } else {
                y[k] = let("mk(" + C + ".x - " + S + ".y, " + C + ".y + " + S + ".x)");
                y[R - k] = let("mk(" + C + ".x + " + S + ".y, " + C + ".y - " + S + ".x)");
            }
        }
        return y;
    }

    /// R = A B: A-point DFTs over x[B i1 + i2], twiddle w_R^{dir i2 k1}, B-point DFTs,
    /// output index k1 + A k2.
    std::vector<std::string> split(std::vector<std::string> const &x, int A, int B) {
        int const R = A * B;
        std::vector<std::vector<std::string>> z(B);
        for (int i2 = 0; i2 < B; ++i2) {
            std::vector<std::string> in(A);
            for (int i1 = 0; i1 < A; ++i1) {
                in[i1] = x[B * i1 + i2];
            }
            z[i2] = dft(in);
            for (int k1 = 0; k1 < A; ++k1) {
                z[i2][k1] = mul_root(z[i2][k1], static_cast<long>(dir_) * i2 * k1, R);
            }
        }
        std::vector<std::string> y(R);
        for (int k1 = 0; k1 < A; ++k1) {
            std::vector<std::string> in(B);
            for (int i2 = 0; i2 < B; ++i2) {
                in[i2] = z[i2][k1];
            }
            auto out = dft(in);
            for (int k2 = 0; k2 < B; ++k2) {
                y[k1 + A * k2] = out[k2];
            }
        }
        return y;
    }

    writer &w_;
    fp_helper f_;
    int dir_;
};

std::string u(std::size_t v) { return std::to_string(v) + "u"; }
std::string ul(std::size_t v) { return std::to_string(v) + "ul"; }

/// Emits one kernel.
class kernel_gen {
  public:
    /// alias_in / alias_out: the input / output pointer of a tile body points into the
    /// work-group's LDS scratch (fused multi-dimensional kernels), so reads of the input and
    /// writes of the output are separated from scratch writes / reads by barriers.
    kernel_gen(fft1d_configuration const &c, std::string name, std::size_t group_size = 0,
               bool alias_in = false, bool alias_out = false, body_options const &opt = {})
        : c_(c), f_{c.fp}, name_(std::move(name)), group_size_(group_size), alias_in_(alias_in),
          alias_out_(alias_out), lds_tables_(opt.lds_tables), lds_barriers_(opt.lds_barriers),
          extra_param_(opt.extra_param), after_input_(opt.after_input),
          input_index_xor_(opt.input_index_xor), single_tile_(opt.single_tile),
          lds_table_pad_(opt.lds_tables && opt.lds_table_pad) {
        nfft_ = c.N_fft();
        L_ = c.factorization.size();
        real_ = c.type != transform_type::c2c;
        half_ = real_ && c.N % 2 == 0;
        pairs_ = real_ && c.N % 2 == 1;
        direct_ = c.direct_radix();
        use_lds_ = L_ > 1 || real_ || direct_;
        MG_ = (c.M + c.Mb - 1) / c.Mb;
        if (c.lds_pad) {
            pad_shift_ = 0;
            while ((std::size_t(1) << pad_shift_) < c.lds_pad) {
                ++pad_shift_;
            }
        }
        np_ = (c.lds_pad ? (nfft_ - 1) + (nfft_ - 1) / c.lds_pad : nfft_ - 1) + 1;
        // 32-bit offsets within a tile when the tile's extent fits comfortably; callbacks,
        // which take absolute offsets, get the tile's uniform 64-bit base plus the 32-bit
        // offset within the tile
        std::size_t const kspan = (pairs_ ? 2 : 1) * c.Kb + 1;
        std::size_t const nspan = c.N + 2;
        std::size_t const ext_in = c.Mb + nspan * c.istride[1] + kspan * c.istride[2];
        std::size_t const ext_out = c.Mb + nspan * c.ostride[1] + kspan * c.ostride[2];
        std::size_t const eb = 2 * static_cast<std::size_t>(c.fp); // largest element size
        local32_ =
            ext_in * eb < (std::size_t(1) << 31) && ext_out * eb < (std::size_t(1) << 31);
        half_in_ = (c.flags & (fft1d_input_f16 | fft1d_input_bf16)) != 0;
        bf16_ = (c.flags & fft1d_input_bf16) != 0;
        mfma_ = (c.flags & fft1d_mfma) != 0;
        prefetch_ = (c.flags & fft1d_prefetch) != 0 && local32_ && !mfma_ &&
                    c.type != transform_type::c2r && !direct_;
    }

    std::string generate() {
        if (half_in_) {
            w_.line(half_prelude());
        }
        emit_tables();
        emit_kernel();
        return w_.str();
    }

    /// Device function computing one tile: `NAME(in, out, K, tile, lds_buf)`, for the
    /// persistent ragged kernel. Work-items beyond Mb * Nt * Kb of a `group_size` work-group
    /// take part in the barriers only.
    std::string generate_body() {
        if (direct_) {
            throw bad_configuration("Tile bodies (ragged and fused kernels) support register "
                                    "radices only");
        }
        prefetch_ = false; // one tile per call: nothing to pipeline
        emit_tables();
        emit_body();
        return w_.str();
    }

  private:
    // ---- helpers -------------------------------------------------------------------------
    std::string T() const { return f_.real(); }
    std::string T2() const { return f_.cplx(); }

    /// Padded LDS index of FFT index n (expression).
    std::string padded(std::string const &n) const {
        return c_.lds_pad ? "(" + n + " + ((" + n + ") >> " + std::to_string(pad_shift_) + "))"
                          : "(" + n + ")";
    }
    /// LDS slot at padded index `pbase` (a variable holding a padded index) plus a constant
    /// padded offset: the offset folds into the LDS instruction's immediate.
    std::string lds_slot(std::string const &pbase, std::size_t offset) const {
        return "col[(" + pbase + " + " + u(offset) + ") * " + u(c_.Mb) + "]";
    }

    /// LDS slot of FFT index n (expression) within this work-item's column.
    std::string lds(std::string const &n) const {
        std::string p = c_.lds_pad ? "(" + n + " + ((" + n + ") >> " +
                                         std::to_string(pad_shift_) + "))"
                                   : "(" + n + ")";
        return "col[" + p + " * " + u(c_.Mb) + "]";
    }

    std::size_t in_elem_bytes() const {
        std::size_t const rb = static_cast<std::size_t>(c_.fp);
        return half_in_ ? 4 : c_.type == transform_type::r2c ? rb : 2 * rb;
    }
    std::size_t out_elem_bytes() const {
        std::size_t const rb = static_cast<std::size_t>(c_.fp);
        return c_.type == transform_type::c2r ? rb : 2 * rb;
    }

    std::string in_type() const {
        if (half_in_) {
            return "unsigned"; // (re, im) half-precision pair
        }
        return c_.type == transform_type::r2c ? T() : T2();
    }
    std::string out_type() const {
        return c_.type == transform_type::c2r ? T() : T2();
    }

    /// Global load of input element at offset expression `off` (elements from `in`).
    std::string gload(std::string const &off) const {
        if (c_.load_function) {
            std::string const abs = local32_ ? "ibase + (unsigned long)(" + off + ")" : off;
            return std::string(c_.load_function) + "((" + in_type() + "*)in, " + abs + ")";
        }
        if (c_.flags & fft1d_nontemporal_load) {
            return "ld_nt(tin + " + off + ")";
        }
        return "tin[" + off + "]";
    }
    /// Global store of `val` at offset `off` (elements from `out`).
    std::string gstore(std::string const &off, std::string const &val) const {
        if (c_.store_function) {
            std::string const abs = local32_ ? "obase + (unsigned long)(" + off + ")" : off;
            return std::string(c_.store_function) + "(out, " + abs + ", " + val + ");";
        }
        if (c_.flags & fft1d_nontemporal_store) {
            return "st_nt(tout + " + off + ", " + val + ");";
        }
        return "tout[" + off + "] = " + val + ";";
    }

    /// Work-group barrier: __syncthreads(), or in fused kernels with lds_barriers a barrier
    /// that orders LDS accesses only (BBFFT_LDS_SYNC, defined by the fused kernel's source).
    std::string sync() const { return lds_barriers_ ? "BBFFT_LDS_SYNC();" : "__syncthreads();"; }
    /// Twiddle tables: __constant__ arrays, or their LDS copies (lds_tables_).
    std::string tw_name(std::size_t s) const {
        return name_ + "_tw" + std::to_string(s) + (lds_tables_ ? "_l" : "");
    }
    std::string twr_name() const { return name_ + "_twr" + (lds_tables_ ? "_l" : ""); }
    /// Entries between the rows (j mod p) of a stage's twiddle table of radix R.
    std::size_t tw_row(std::size_t R) const { return R + (lds_table_pad_ ? 1 : 0); }
    std::string dtab_name() const { return name_ + "_dtab"; }

    void emit_table(std::string const &name, std::vector<std::pair<long, long>> const &roots,
                    bool times_i) {
        w_.line("static __constant__ " + T2() + " " + name + "[" + std::to_string(roots.size()) +
                "] = {");
        std::string row;
        for (std::size_t i = 0; i < roots.size(); ++i) {
            auto [re, im] = f_.root(roots[i].first, roots[i].second);
            if (times_i) {
                // i * (re + i im) = -im + i re
                std::string nim = im[0] == '-' ? im.substr(1) : "-" + im;
                row += "{" + nim + ", " + re + "}, ";
            } else {
                row += "{" + re + ", " + im + "}, ";
            }
            if ((i + 1) % 4 == 0 || i + 1 == roots.size()) {
                w_.line("  " + row);
                row.clear();
            }
        }
        w_.line("};");
    }

    std::string mfma_table_name() const { return name_ + "_mA"; }

    /// Terms of the 32 x 32 real matrix of the radix-16 DFT acting on interleaved (re, im)
    /// vectors: row 2 k1 + co, column 2 n1 + ci, bit patterns in the input's half format.
    void emit_mfma_table() {
        unsigned const T = mfma_split_terms(c_.flags);
        std::vector<std::uint16_t> bits(T * 1024u);
        for (unsigned i = 0; i < 32; ++i) {
            for (unsigned q = 0; q < 32; ++q) {
                unsigned const k1 = i / 2, co = i % 2, n1 = q / 2, ci = q % 2;
                auto w = gen::root_of_unity(static_cast<long>(c_.direction) * long(n1 * k1), 16);
                double a = co == 0 ? (ci == 0 ? w.real() : -w.imag())
                                   : (ci == 0 ? w.imag() : w.real());
                for (unsigned t = 0; t < T; ++t) {
                    double h = bf16_ ? round_bf16(a) : round_f16(a);
                    bits[t * 1024u + i * 32u + q] = bf16_ ? bf16_bits(h) : f16_bits(h);
                    a -= h; // exact: h is a rounding of a
                }
            }
        }
        w_.line("static __constant__ __attribute__((aligned(16))) unsigned short " +
                mfma_table_name() + "[" + std::to_string(bits.size()) + "] = {");
        std::string row;
        for (std::size_t i = 0; i < bits.size(); ++i) {
            row += std::to_string(bits[i]) + "u, ";
            if ((i + 1) % 16 == 0) {
                w_.line("  " + row);
                row.clear();
            }
        }
        w_.line("};");
    }

    void emit_tables() {
        if (mfma_) {
            emit_mfma_table();
        }
        std::size_t p = c_.factorization[0];
        for (std::size_t s = 1; s < L_; ++s) {
            std::size_t R = c_.factorization[s];
            std::vector<std::pair<long, long>> roots;
            for (std::size_t jp = 0; jp < p; ++jp) {
                for (std::size_t i = 0; i < R; ++i) {
                    roots.push_back({static_cast<long>(c_.direction) * long(i) * long(jp),
                                     long(p * R)});
                }
            }
            emit_table(name_ + "_tw" + std::to_string(s), roots, false);
            p *= R;
        }
        if (direct_) {
            // (cos, sin)(2 pi m / R), m = 0..R-1, of the direct stage
            std::vector<std::pair<long, long>> roots;
            for (long m = 0; m < direct_; ++m) {
                roots.push_back({m, direct_});
            }
            emit_table(dtab_name(), roots, false);
        }
        if (half_) {
            // i * w_N^{dir q}, q = 0..H/2 (reference: sbfft_gen.cpp:192,238)
            std::vector<std::pair<long, long>> roots;
            for (std::size_t q = 0; q <= nfft_ / 2; ++q) {
                roots.push_back({static_cast<long>(c_.direction) * long(q), long(c_.N)});
            }
            emit_table(name_ + "_twr", roots, true);
        }
        if (lds_tables_) {
            emit_lds_tables();
        }
    }

    /// LDS copies of the twiddle tables and `NAME_tables(lid, n)`, which fills them with the
    /// n work-items lid = 0..n-1 (the caller orders it before the first use with a barrier).
    void emit_lds_tables() {
        struct table {
            std::string name;
            std::size_t entries, row; // row: entries per row of the constant table (0: one row)
        };
        std::vector<table> tabs;
        std::size_t p = c_.factorization[0];
        for (std::size_t s = 1; s < L_; ++s) {
            tabs.push_back({name_ + "_tw" + std::to_string(s), p * c_.factorization[s],
                            c_.factorization[s]});
            p *= c_.factorization[s];
        }
        if (half_) {
            tabs.push_back({name_ + "_twr", nfft_ / 2 + 1, 0});
        }
        for (auto const &t : tabs) {
            std::size_t const n = t.row ? t.entries / t.row * tw_row(t.row) : t.entries;
            w_.line("__shared__ " + T2() + " " + t.name + "_l[" + std::to_string(n) + "];");
        }
        w_.line("static __device__ __forceinline__ void " + name_ +
                "_tables(unsigned lid, unsigned n) {");
        w_.open("");
        for (auto const &t : tabs) {
            std::string dst = "q";
            if (t.row && tw_row(t.row) != t.row) {
                dst = "(q / " + u(t.row) + ") * " + u(tw_row(t.row)) + " + q % " + u(t.row);
            }
            w_.line("for (unsigned q = lid; q < " + u(t.entries) + "; q += n) " + t.name + "_l[" +
                    dst + "] = " + t.name + "[q];");
        }
        w_.close();
        w_.line("}");
    }

    // ---- kernel ----------------------------------------------------------------------------
    void emit_kernel() {
        std::size_t const wg = c_.work_group_size();
        std::string const occupancy =
            c_.waves_per_eu ? " __attribute__((amdgpu_waves_per_eu(" +
                                  std::to_string(c_.waves_per_eu) + ")))"
                            : "";
        w_.line("extern \"C\" __global__ void __launch_bounds__(" + std::to_string(wg) + ")" +
                occupancy + " " +
                name_ + "(const " + in_type() + "* in, " + out_type() +
                "* out, unsigned long K) {");
        w_.open("");
        if (use_lds_) {
            w_.line("__shared__ " + T2() + " lds_buf[" + std::to_string(c_.Mb * c_.Kb * np_) +
                    "];");
        }
        w_.line("const unsigned lid = threadIdx.x;");
        if (direct_) {
            // roots of the direct stage in LDS (read with per-lane indices); the first tile's
            // barrier after the direct stage's input phase orders these writes before any read
            w_.line("__shared__ " + T2() + " dts[" + std::to_string(direct_) + "];");
            w_.line("for (unsigned q = lid; q < " + u(direct_) + "; q += " + u(wg) + ") dts[q] = " +
                    dtab_name() + "[q];");
        }
        w_.line("const unsigned ml = lid % " + u(c_.Mb) + ";");
        w_.line("const unsigned t = (lid / " + u(c_.Mb) + ") % " + u(c_.Nt) + ";");
        w_.line("const unsigned kl = lid / " + u(c_.Mb * c_.Nt) + ";");
        if (pairs_) {
            w_.line("const unsigned long kcols = (K + 1ul) / 2ul;");
        } else {
            w_.line("const unsigned long kcols = K;");
        }
        w_.line("const unsigned long ntiles = " + ul(MG_) + " * ((kcols + " + ul(c_.Kb - 1) +
                ") / " + ul(c_.Kb) + ");");
        if (use_lds_) {
            w_.line(T2() + "* col = lds_buf + ml + " + u(c_.Mb * np_) + " * kl;");
        }
        if (prefetch_) {
            // Stage-0 inputs of the work-group's first tile; each iteration then loads those
            // of the next tile while it transforms the current one.
            for (auto const &q : prefetch_regs()) {
                for (auto const &qi : q) {
                    w_.line(T2() + " " + qi + ";");
                }
            }
            w_.open("if (blockIdx.x < ntiles)");
            w_.line("const unsigned long tile = blockIdx.x;");
            emit_tile_origin();
            emit_stage0_loads();
            w_.close();
        }
        w_.open("for (unsigned long tile = blockIdx.x; tile < ntiles; tile += gridDim.x)");
        emit_tile();
        if (use_lds_) {
            w_.line(sync());
        }
        w_.close();
        w_.close();
        w_.line("}");
    }

    void emit_body() {
        w_.line("static __device__ __forceinline__ void " + name_ + "(const " + in_type() +
                "* in, " + out_type() + "* out, unsigned long K, unsigned long tile, " + T2() +
                "* lds_buf, unsigned lid" + (extra_param_.empty() ? "" : ", " + extra_param_) +
                ") {");
        w_.open("");
        w_.line("const unsigned ml = lid % " + u(c_.Mb) + ";");
        w_.line("const unsigned t = (lid / " + u(c_.Mb) + ") % " + u(c_.Nt) + ";");
        w_.line("const unsigned kl = lid / " + u(c_.Mb * c_.Nt) + ";");
        if (group_size_ > c_.work_group_size()) {
            band_ = "inb";
            w_.line("const bool inb = kl < " + u(c_.Kb) + ";");
        }
        if (pairs_) {
            w_.line("const unsigned long kcols = (K + 1ul) / 2ul;");
        } else {
            w_.line("const unsigned long kcols = K;");
        }
        if (use_lds_) {
            std::string klc = band_.empty() ? "kl" : "(inb ? kl : 0u)";
            w_.line(T2() + "* col = lds_buf + ml + " + u(c_.Mb * np_) + " * " + klc + ";");
        }
        emit_tile();
        w_.close();
        w_.line("}");
    }

    /// One (M tile, K tile) of work: `tile` selects it.
    void emit_tile() {
        emit_tile_origin();
        if (c_.type == transform_type::c2r) {
            emit_c2r_pre();
            w_.line(sync());
        }
        for (std::size_t s = 0; s < L_; ++s) {
            if (s == 0 && mfma_) {
                emit_mfma_stage0();
            } else if (is_direct_radix(c_.factorization[s])) {
                emit_direct_stage(s);
            } else {
                emit_stage(s);
            }
            if (s == 0 && !after_input_.empty()) {
                // stage 0 ended with a barrier: every work-item's input reads are complete
                if (L_ == 1 && c_.type != transform_type::r2c) {
                    throw bad_configuration("after-input code needs a barrier after stage 0");
                }
                w_.line(after_input_);
            }
        }
        if (c_.type == transform_type::r2c) {
            emit_r2c_post();
        }
    }

    /// Register names q<b>_<i> holding the prefetched stage-0 inputs (butterfly b, input i).
    std::vector<std::vector<std::string>> prefetch_regs() const {
        std::size_t const R = c_.factorization[0];
        std::size_t const NR = nfft_ / R;
        std::size_t const B = (NR + c_.Nt - 1) / c_.Nt;
        std::vector<std::vector<std::string>> q(B, std::vector<std::string>(R));
        for (std::size_t b = 0; b < B; ++b) {
            for (std::size_t i = 0; i < R; ++i) {
                q[b][i] = "q" + std::to_string(b) + "_" + std::to_string(i);
            }
        }
        return q;
    }

    /// Loads the stage-0 inputs of the tile whose origin is in scope into the prefetch
    /// registers (zeros for work-items without a butterfly or column).
    void emit_stage0_loads() {
        std::size_t const R = c_.factorization[0];
        std::size_t const NR = nfft_ / R;
        auto const q = prefetch_regs();
        for (std::size_t b = 0; b < q.size(); ++b) {
            std::string const j = "j0_" + std::to_string(b);
            w_.line("const unsigned " + j + " = t + " + u(b * c_.Nt) + ";");
            std::string cond = "active";
            if ((b + 1) * c_.Nt > NR) {
                cond += " && " + j + " < " + u(NR);
            }
            w_.open("if (" + cond + ")");
            for (std::size_t i = 0; i < R; ++i) {
                emit_global_input(q[b][i], j + " + " + u(i * NR));
            }
            w_.close();
            w_.open("else");
            for (std::size_t i = 0; i < R; ++i) {
                w_.line(q[b][i] + " = " + f_.zero2() + ";");
            }
            w_.close();
        }
    }

    /// Origin, activity and base pointers of the tile `tile` (in scope).
    void emit_tile_origin() {
        // tile origin (uniform across the work-group)
        if (MG_ == 1) {
            w_.line("const unsigned long m0 = 0ul;");
            w_.line("const unsigned long k0 = tile * " + ul(c_.Kb) + ";");
        } else {
            w_.line("const unsigned long m0 = (tile % " + ul(MG_) + ") * " + ul(c_.Mb) + ";");
            w_.line("const unsigned long k0 = (tile / " + ul(MG_) + ") * " + ul(c_.Kb) + ";");
        }
        w_.line("const unsigned long m = m0 + ml;");
        w_.line("const unsigned long kc = k0 + kl;");
        std::string act = "kc < kcols";
        if (MG_ * c_.Mb != c_.M) {
            act = "m < " + ul(c_.M) + " && " + act;
        }
        if (single_tile_) {
            if (MG_ * c_.Mb != c_.M) {
                throw bad_configuration("a single-tile body needs Mb = M");
            }
            act = band_.empty() ? "true" : band_;
        } else if (!band_.empty()) {
            act = band_ + " && " + act;
        }
        w_.line("const bool active = " + act + ";");
        std::string const kf = pairs_ ? "2ul * " : "";
        if (pairs_) {
            w_.line("const bool active_b = active && 2ul * kc + 1ul < K;");
        }
        if (local32_) {
            // Uniform tile base pointers and 32-bit offsets within the tile: the per-lane
            // address arithmetic stays 32-bit and the base can live in scalar registers.
            w_.line("const " + in_type() + "* tin = in + m0 + " + kf + "k0 * " +
                    ul(c_.istride[2]) + ";");
            w_.line(out_type() + "* tout = out + m0 + " + kf + "k0 * " + ul(c_.ostride[2]) +
                    ";");
            std::string const kl2 = pairs_ ? "2u * kl" : "kl";
            w_.line("const unsigned ib = ml + " + kl2 + " * " + u(c_.istride[2]) + ";");
            w_.line("const unsigned ob = ml + " + kl2 + " * " + u(c_.ostride[2]) + ";");
            if (c_.load_function) {
                w_.line("const unsigned long ibase = m0 + " + kf + "k0 * " +
                        ul(c_.istride[2]) + "; // tile origin (callback offsets)");
            }
            if (c_.store_function) {
                w_.line("const unsigned long obase = m0 + " + kf + "k0 * " +
                        ul(c_.ostride[2]) + "; // tile origin (callback offsets)");
            }
        } else {
            w_.line("const " + in_type() + "* tin = in;");
            w_.line(out_type() + "* tout = out;");
            w_.line("const unsigned long ib = m + " + kf + "kc * " + ul(c_.istride[2]) + ";");
            w_.line("const unsigned long ob = m + " + kf + "kc * " + ul(c_.ostride[2]) + ";");
        }
    }

    /// LDS slot of FFT index n (expression) in the column that starts at pointer `p`.
    std::string lds_at(std::string const &p, std::string const &n) const {
        std::string q = c_.lds_pad ? "(" + n + " + ((" + n + ") >> " +
                                         std::to_string(pad_shift_) + "))"
                                   : "(" + n + ")";
        return p + "[" + q + " * " + u(c_.Mb) + "]";
    }

    /**
     * Stage 0 (radix 16, p = 1) as a GEMM on the matrix cores. The work-group tile holds
     * Mb * (N / 16) * Kb columns (m, j, k) of 16 complex inputs x[j + (N / 16) n1]; each wave
     * takes 32 columns at a time: B (32 x 32, inner index 2 n1 + re/im) is loaded straight from
     * global memory in the input's half format, A is the split DFT matrix, and
     * D = sum over terms of A_t B (fp32 accumulation, smallest term first) holds the 16 outputs
     * of each column, written to LDS slot j * 16 + k1 like the vector stage 0.
     *
     * Operand layouts of v_mfma_f32_32x32x16_{f16,bf16} (64 lanes): A row i and B column j
     * are lane % 32; each lane holds 8 consecutive inner indices, lanes 32..63 the second 8
     * (any inner-index assignment shared by A and B gives the same product); D element r of
     * lane l is row 8 (r / 4) + 4 (l / 32) + r % 4 of column l % 32.
     */
    void emit_mfma_stage0() {
        std::size_t const N2 = nfft_ / 16;
        std::size_t const ncols = c_.Mb * N2 * c_.Kb;
        std::size_t const nwaves = (c_.work_group_size() + 63) / 64;
        std::size_t const ntile = (ncols + 31) / 32;
        std::size_t const iters = (ntile + nwaves - 1) / nwaves;
        unsigned const T = mfma_split_terms(c_.flags);
        std::string const vt = bf16_ ? "bbfft_b8" : "bbfft_h8";
        std::string const op = bf16_ ? "__builtin_amdgcn_mfma_f32_32x32x16_bf16"
                                     : "__builtin_amdgcn_mfma_f32_32x32x16_f16";
        w_.line("// stage 0: radix 16 as a DFT-matrix GEMM on MFMA (" + std::to_string(T) +
                "-term split matrix, fp32 accumulation)");
        w_.open("");
        w_.line("const unsigned lane = lid % 64u, wv = lid / 64u;");
        w_.line("const unsigned hh = lane / 32u, cc = lane % 32u;");
        for (unsigned t = 0; t < T; ++t) {
            for (unsigned h2 = 0; h2 < 2; ++h2) {
                w_.line("const " + vt + " a" + std::to_string(t) + "_" + std::to_string(h2) +
                        " = __builtin_bit_cast(" + vt + ", *(const bbfft_u4*)(" +
                        mfma_table_name() + " + " + u(t * 1024 + h2 * 16) +
                        " + cc * 32u + hh * 8u));");
            }
        }
        // All loads of the work-item are issued before the matrix instructions (one memory
        // latency per tile instead of one per 32 columns).
        std::string ld_cond = "cv && k0 + kk < kcols";
        if (MG_ * c_.Mb != c_.M) {
            ld_cond += " && m0 + mm < " + ul(c_.M);
        }
        for (std::size_t it = 0; it < iters; ++it) {
            std::string const x = "_" + std::to_string(it);
            w_.line("const unsigned cl" + x + " = " + u((it * nwaves) * 32) +
                    " + wv * 32u + cc;");
            w_.line("const bool cv" + x + " = cl" + x + " < " + u(ncols) + ";");
            w_.line("const unsigned mm" + x + " = cl" + x + " % " + u(c_.Mb) + ";");
            w_.line("const unsigned jj" + x + " = (cl" + x + " / " + u(c_.Mb) + ") % " + u(N2) +
                    ";");
            w_.line("const unsigned kk" + x + " = cl" + x + " / " + u(c_.Mb * N2) + ";");
            std::string ldx = ld_cond;
            for (std::string v : {"cv", "kk", "mm"}) {
                std::size_t pos = 0;
                while ((pos = ldx.find(v, pos)) != std::string::npos) {
                    ldx.insert(pos + v.size(), x);
                    pos += v.size() + x.size();
                }
            }
            w_.line("const bool ld" + x + " = " + ldx + ";");
            if (local32_) {
                w_.line("const unsigned bb" + x + " = mm" + x + " + kk" + x + " * " +
                        u(c_.istride[2]) + ";");
            } else {
                w_.line("const unsigned long bb" + x + " = m0 + mm" + x + " + (k0 + kk" + x +
                        ") * " + ul(c_.istride[2]) + ";");
            }
            for (unsigned h2 = 0; h2 < 2; ++h2) {
                std::string r = "r" + std::to_string(h2) + x;
                w_.line("bbfft_u4 " + r + " = {0u, 0u, 0u, 0u};");
                w_.open("if (ld" + x + ")");
                for (unsigned e = 0; e < 4; ++e) {
                    std::string n1 = "(" + u(8 * h2 + e) + " + 4u * hh)";
                    std::string off =
                        local32_ ? "bb" + x + " + (jj" + x + " + " + u(N2) + " * " + n1 + ") * " +
                                       u(c_.istride[1])
                                 : "bb" + x + " + (unsigned long)(jj" + x + " + " + u(N2) +
                                       " * " + n1 + ") * " + ul(c_.istride[1]);
                    w_.line(r + "[" + std::to_string(e) + "] = " + gload(off) + ";");
                }
                w_.close();
            }
        }
        for (std::size_t it = 0; it < iters; ++it) {
            std::string const x = "_" + std::to_string(it);
            w_.open("");
            w_.line("bbfft_f16v acc = {};");
            for (unsigned t = T; t-- > 0;) {
                for (unsigned h2 = 0; h2 < 2; ++h2) {
                    w_.line("acc = " + op + "(a" + std::to_string(t) + "_" + std::to_string(h2) +
                            ", __builtin_bit_cast(" + vt + ", r" + std::to_string(h2) + x +
                            "), acc, 0, 0, 0);");
                }
            }
            w_.open("if (cv" + x + ")");
            w_.line(T2() + "* cp = lds_buf + mm" + x + " + " + u(c_.Mb * np_) + " * kk" + x +
                    ";");
            // slot jj * 16 + k1 (k1 < 16): with no padding or a padding interval of at least
            // 16, the padded index is padded(jj * 16) + k1
            bool const split = c_.lds_pad == 0 || c_.lds_pad >= 16;
            if (split) {
                w_.line("const unsigned pj = " + padded("jj" + x + " * 16u") + " + 2u * hh;");
            }
            for (unsigned r = 0; r < 16; r += 2) {
                std::size_t const kc = 4 * (r / 4) + (r % 4) / 2;
                std::string const val = "mk(acc[" + std::to_string(r) + "], acc[" +
                                        std::to_string(r + 1) + "])";
                if (split) {
                    w_.line("cp[(pj + " + u(kc) + ") * " + u(c_.Mb) + "] = " + val + ";");
                } else {
                    std::string k1 = "(" + u(kc) + " + 2u * hh)";
                    w_.line(lds_at("cp", "jj" + x + " * 16u + " + k1) + " = " + val + ";");
                }
            }
            w_.close();
            w_.close();
        }
        w_.close();
        w_.line(sync());
    }

    /// Offset (elements) of FFT index n in the input column along mode 1.
    std::string ioff(std::string const &n, std::size_t extra_k = 0) const {
        if (local32_) {
            std::string k = extra_k ? " + " + u(c_.istride[2] * extra_k) : "";
            return "ib" + k + " + (unsigned)(" + n + ") * " + u(c_.istride[1]);
        }
        std::string k = extra_k ? " + " + ul(c_.istride[2] * extra_k) : "";
        return "ib" + k + " + (unsigned long)(" + n + ") * " + ul(c_.istride[1]);
    }
    std::string ooff(std::string const &n, std::size_t extra_k = 0) const {
        if (local32_) {
            std::string k = extra_k ? " + " + u(c_.ostride[2] * extra_k) : "";
            return "ob" + k + " + (unsigned)(" + n + ") * " + u(c_.ostride[1]);
        }
        std::string k = extra_k ? " + " + ul(c_.ostride[2] * extra_k) : "";
        return "ob" + k + " + (unsigned long)(" + n + ") * " + ul(c_.ostride[1]);
    }

    bool vector_real_in() const {
        return half_ && c_.type == transform_type::r2c && !c_.load_function &&
               c_.istride[1] == 1 && c_.istride[2] % 2 == 0 && c_.M == 1;
    }
    bool vector_real_out() const {
        return half_ && c_.type == transform_type::c2r && !c_.store_function &&
               c_.ostride[1] == 1 && c_.ostride[2] % 2 == 0 && c_.M == 1;
    }

    /// Load z[n] of the complex FFT input from global memory into variable v.
    void emit_global_input(std::string const &v, std::string const &n0) {
        std::string const n =
            input_index_xor_.empty() ? n0 : "((" + n0 + ") ^ " + input_index_xor_ + ")";
        if (!input_index_xor_.empty() && !(c_.type == transform_type::c2c && !half_in_) &&
            !(half_ && vector_real_in())) {
            throw bad_configuration("permuted input rows need complex element reads");
        }
        if (c_.type == transform_type::c2c) {
            if (half_in_) {
                w_.line(v + " = " + std::string(bf16_ ? "bf16c_to_f2(" : "f16c_to_f2(") +
                        gload(ioff(n)) + ");");
            } else {
                w_.line(v + " = " + gload(ioff(n)) + ";");
            }
        } else if (half_) {
            if (vector_real_in()) {
                std::string p = "((const " + T2() + "*)(tin + ib)) + (" + n + ")";
                w_.line(v + " = " + ((c_.flags & fft1d_nontemporal_load) ? "ld_nt(" + p + ")"
                                                                          : "*(" + p + ")") +
                        ";");
            } else {
                w_.line(v + " = mk(" + gload(ioff("2u * (" + n + ")")) + ", " +
                        gload(ioff("2u * (" + n + ") + 1u")) + ");");
            }
        } else { // pairs
            w_.line(v + ".x = " + gload(ioff(n)) + ";");
            w_.line(v + ".y = active_b ? " + gload(ioff(n, 1)) + " : " + f_.lit(0.0) + ";");
        }
    }

    /// Store FFT output y at index n to global memory.
    void emit_global_output(std::string const &y, std::string const &n) {
        if (c_.type == transform_type::c2c) {
            w_.line(gstore(ooff(n), y));
        } else if (half_) {
            if (vector_real_out()) {
                std::string p = "((" + T2() + "*)(tout + ob)) + (" + n + ")";
                w_.line((c_.flags & fft1d_nontemporal_store) ? "st_nt(" + p + ", " + y + ");"
                                                             : "*(" + p + ") = " + y + ";");
            } else {
                w_.line(gstore(ooff("2u * (" + n + ")"), y + ".x"));
                w_.line(gstore(ooff("2u * (" + n + ") + 1u"), y + ".y"));
            }
        } else { // pairs
            w_.line(gstore(ooff(n), y + ".x"));
            w_.line("if (active_b) " + gstore(ooff(n, 1), y + ".y"));
        }
    }

    void emit_stage(std::size_t s) {
        std::size_t const R = c_.factorization[s];
        std::size_t p = 1;
        for (std::size_t i = 0; i < s; ++i) {
            p *= c_.factorization[i];
        }
        std::size_t const NR = nfft_ / R;
        std::size_t const B = (NR + c_.Nt - 1) / c_.Nt;
        bool const src_global = s == 0 && c_.type != transform_type::c2r;
        bool const dst_global = s + 1 == L_ && c_.type != transform_type::r2c;
        // Reads and writes of LDS (or of memory aliasing it) are separated by a barrier.
        bool const reads_shared = !src_global || alias_in_;
        bool const writes_shared = !dst_global || alias_out_;
        bool const mid_barrier = reads_shared && writes_shared;

        w_.line("// stage " + std::to_string(s) + ": radix " + std::to_string(R) + ", p = " +
                std::to_string(p));
        std::vector<std::vector<std::string>> v(B);
        std::vector<std::string> guard(B), jn(B);
        for (std::size_t b = 0; b < B; ++b) {
            jn[b] = "j" + std::to_string(s) + "_" + std::to_string(b);
            w_.line("const unsigned " + jn[b] + " = t + " + u(b * c_.Nt) + ";");
            if ((b + 1) * c_.Nt > NR) {
                guard[b] = jn[b] + " < " + u(NR);
            }
            v[b].resize(R);
            for (std::size_t i = 0; i < R; ++i) {
                v[b][i] = "v" + std::to_string(s) + "_" + std::to_string(b) + "_" +
                          std::to_string(i);
                w_.line(T2() + " " + v[b][i] + ";");
            }
            if (src_global && prefetch_) {
                auto const q = prefetch_regs();
                for (std::size_t i = 0; i < R; ++i) {
                    w_.line(v[b][i] + " = " + q[b][i] + ";");
                }
                continue;
            }
            // Condition of the butterfly's computation (below) and of its loads: work-items
            // that compute without loading (no column, inside the band) hold zeros.
            std::string ccond = guard[b];
            if (!band_.empty()) {
                ccond = ccond.empty() ? band_ : band_ + " && " + ccond;
            }
            std::string cond = ccond;
            if (src_global && !single_tile_) {
                cond = guard[b].empty() ? "active" : "active && " + guard[b];
            }
            bool const zero_fill = cond != ccond;
            if (!cond.empty()) {
                w_.open("if (" + cond + ")");
            }
            // Reads at j + i NR: with NR a multiple of the padding interval, the padded index
            // is padded(j) + i (NR + NR / pad), one base per butterfly.
            std::size_t const padv = c_.lds_pad;
            bool const rd_split = !src_global && (padv == 0 || NR % padv == 0);
            std::string const rb = "pr" + std::to_string(s) + "_" + std::to_string(b);
            if (rd_split) {
                w_.line("const unsigned " + rb + " = " + padded(jn[b]) + ";");
            }
            for (std::size_t i = 0; i < R; ++i) {
                std::string n = jn[b] + " + " + u(i * NR);
                if (src_global) {
                    emit_global_input(v[b][i], n);
                } else if (rd_split) {
                    w_.line(v[b][i] + " = " +
                            lds_slot(rb, i * (NR + (padv ? NR / padv : 0))) + ";");
                } else {
                    w_.line(v[b][i] + " = " + lds(n) + ";");
                }
            }
            if (s > 0) {
                std::string jp = "(" + jn[b] + " % " + u(p) + ") * " + u(tw_row(R));
                for (std::size_t i = 1; i < R; ++i) {
                    w_.line(v[b][i] + " = c_mul(" + v[b][i] + ", " + tw_name(s) + "[" + jp +
                            " + " + u(i) + "]);");
                }
            }
            if (!cond.empty()) {
                w_.close();
                if (zero_fill) {
                    // the branch is skipped entirely when a whole wave is active
                    w_.open("else");
                    for (std::size_t i = 0; i < R; ++i) {
                        w_.line(v[b][i] + " = " + f_.zero2() + ";");
                    }
                    w_.close();
                }
            }
        }
        if (src_global && prefetch_) {
            // the next tile's inputs are in flight while this tile is transformed
            w_.open("");
            w_.line("const unsigned long tile_next = tile + gridDim.x;");
            w_.open("if (tile_next < ntiles)");
            w_.line("const unsigned long tile = tile_next;");
            emit_tile_origin();
            emit_stage0_loads();
            w_.close();
            w_.close();
        }
        if (mid_barrier) {
            w_.line(sync());
        }
        for (std::size_t b = 0; b < B; ++b) {
            std::string cond = guard[b];
            if (dst_global) {
                cond = cond.empty() ? "active" : "active && " + cond;
            }
            // The butterfly is computed unconditionally when only the store is predicated;
            // scoping keeps the SSA temporaries local.
            std::string ccond = guard[b];
            if (!band_.empty()) {
                ccond = ccond.empty() ? band_ : band_ + " && " + ccond;
            }
            if (!ccond.empty()) {
                w_.open("if (" + ccond + ")");
            } else {
                w_.open("");
            }
            dft_gen d(w_, f_, c_.direction);
            auto y = d.dft(v[b]);
            std::string const &j = jn[b];
            if (dst_global) {
                if (guard[b].empty()) {
                    w_.open("if (active)");
                } else {
                    w_.open("if (active)");
                }
                for (std::size_t i = 0; i < R; ++i) {
                    emit_global_output(y[i], j + " + " + u(i * NR));
                }
                w_.close();
            } else {
                std::string base = p == 1 ? j + " * " + u(R)
                                          : "(" + j + " / " + u(p) + ") * " + u(p * R) + " + " +
                                                j + " % " + u(p);
                // Writes at base + i p: split into padded(base) + a constant when no padding
                // carry depends on the work-item (p a multiple of the interval, or p = 1 with
                // R a multiple or a power-of-two divisor of it).
                std::size_t const padv = c_.lds_pad;
                auto is_pow2 = [](std::size_t v) { return v && !(v & (v - 1)); };
                int mode = -1; // -1: no split
                if (padv == 0 || p % padv == 0) {
                    mode = 0;
                } else if (p == 1 && R % padv == 0) {
                    mode = 1;
                } else if (p == 1 && is_pow2(R) && padv % R == 0) {
                    mode = 2;
                }
                std::string const wb = "pw" + std::to_string(s) + "_" + std::to_string(b);
                if (mode >= 0) {
                    w_.line("const unsigned " + wb + " = " + padded(base) + ";");
                }
                for (std::size_t i = 0; i < R; ++i) {
                    std::size_t off = 0;
                    switch (mode) {
                    case 0:
                        off = i * p + (padv ? i * p / padv : 0);
                        break;
                    case 1:
                        off = i + i / padv;
                        break;
                    case 2:
                        off = i;
                        break;
                    default:
                        break;
                    }
                    if (mode >= 0) {
                        w_.line(lds_slot(wb, off) + " = " + y[i] + ";");
                    } else {
                        w_.line(lds(base + " + " + u(i * p)) + " = " + y[i] + ";");
                    }
                }
            }
            w_.close();
        }
        if (!dst_global) {
            w_.line(sync());
        }
    }

    /**
     * Stage s with an odd radix R above the register limit as a direct DFT (the reference's
     * basic_esum fallback for large prime factors, src/base/mixed_radix_fft.cpp:46-53, in the
     * symmetric-pair form of its pair optimisation, mixed_radix_fft.cpp:56-118):
     *   y_k = x_0 + sum_i a_i cos(2 pi i k / R) + dir i sum_i b_i sin(2 pi i k / R),
     *   y_{R-k} = the same with the sine sum negated, a_i = x_i + x_{R-i}, b_i = x_i - x_{R-i}.
     * Phase 1: units (j, i), i = 0..h = (R-1)/2, read x_i and x_{R-i} of butterfly j (global
     * memory or LDS, twiddled by w_{pR}^{i (j mod p)} for s > 0) and leave x_0, a_i, b_i in the
     * butterfly's LDS slots j + i NR and j + (R - i) NR. Phase 2: work-item t serves butterfly
     * j = t mod NR and the output pairs (k, R - k), k = t / NR + b Tk, accumulating the sums
     * over i in registers (two partial sums in fp32) with the roots read from an LDS table at
     * (i k) mod R; after a barrier the outputs go to LDS (Stockham order) or global memory.
     */
    void emit_direct_stage(std::size_t s) {
        std::size_t const R = c_.factorization[s];
        std::size_t p = 1;
        for (std::size_t i = 0; i < s; ++i) {
            p *= c_.factorization[i];
        }
        std::size_t const NR = nfft_ / R;
        std::size_t const h = (R - 1) / 2;
        std::size_t const U = h + 1;      // output units (k, R - k) per butterfly, k = 0..h
        std::size_t const Tk = c_.Nt / NR; // work-items per butterfly
        std::size_t const Bu = (U + Tk - 1) / Tk;
        std::size_t const parts = c_.fp == precision::f32 ? 2 : 1;
        bool const src_global = s == 0 && c_.type != transform_type::c2r;
        bool const dst_global = s + 1 == L_ && c_.type != transform_type::r2c;
        std::string const S = std::to_string(s);
        w_.line("// stage " + S + ": radix " + std::to_string(R) + " as a direct DFT, p = " +
                std::to_string(p));
        // phase 1
        std::size_t const P1 = NR * U;
        std::size_t const B1 = (P1 + c_.Nt - 1) / c_.Nt;
        for (std::size_t b = 0; b < B1; ++b) {
            std::string const x = S + "_" + std::to_string(b);
            w_.open("");
            w_.line("const unsigned du = t + " + u(b * c_.Nt) + ";");
            std::string cond = (b + 1) * c_.Nt > P1 ? "du < " + u(P1) : "";
            if (cond.empty()) {
                w_.open("");
            } else {
                w_.open("if (" + cond + ")");
            }
            w_.line("const unsigned dj = " + (NR == 1 ? std::string("0u") : "du % " + u(NR)) +
                    ";");
            w_.line("const unsigned di = " + (NR == 1 ? std::string("du") : "du / " + u(NR)) +
                    ";");
            // x_{R-i} of unit i = 0 does not exist: its index falls back to x_0's
            w_.line("const unsigned dn1 = dj + di * " + u(NR) + ";");
            w_.line("const unsigned dn2 = di == 0u ? dj : dj + (" + u(R) + " - di) * " + u(NR) +
                    ";");
            w_.line(T2() + " x1 = " + f_.zero2() + ", x2 = " + f_.zero2() + ";");
            if (src_global) {
                w_.open("if (active)");
                emit_global_input("x1", "dn1");
                emit_global_input("x2", "dn2");
                w_.close();
            } else {
                w_.line("x1 = " + lds("dn1") + ";");
                w_.line("x2 = " + lds("dn2") + ";");
            }
            if (s > 0) {
                std::string jp = "(dj % " + u(p) + ") * " + u(tw_row(R));
                w_.line("x1 = c_mul(x1, " + tw_name(s) + "[" + jp + " + di]);");
                w_.line("x2 = c_mul(x2, " + tw_name(s) + "[" + jp + " + (di == 0u ? 0u : " + u(R) +
                        " - di)]);");
            }
            w_.open("if (di == 0u)");
            w_.line(lds("dn1") + " = x1;");
            w_.close();
            w_.open("else");
            w_.line(lds("dn1") + " = c_add(x1, x2);");
            w_.line(lds("dn2") + " = c_sub(x1, x2);");
            w_.close();
            w_.close();
            w_.close();
        }
        w_.line(sync());
        // phase 2
        w_.open("");
        w_.line("const unsigned dj = " + (NR == 1 ? std::string("0u") : "t % " + u(NR)) + ";");
        w_.line("const unsigned dk = " + (NR == 1 ? std::string("t") : "t / " + u(NR)) + ";");
        w_.line("const " + T2() + " x0 = " + lds("dj") + ";");
        std::vector<std::string> kk(Bu), valid(Bu);
        for (std::size_t b = 0; b < Bu; ++b) {
            std::string const x = std::to_string(b);
            kk[b] = "dkk" + x;
            w_.line("const unsigned dk" + x + " = dk + " + u(b * Tk) + ";");
            if ((b + 1) * Tk > U) {
                valid[b] = "dk" + x + " < " + u(U);
                // invalid units run with k = 0 (in-bounds table reads), their sums are unused
                w_.line("const unsigned " + kk[b] + " = " + valid[b] + " ? dk" + x + " : 0u;");
            } else {
                w_.line("const unsigned " + kk[b] + " = dk" + x + ";");
            }
            w_.line("unsigned dm" + x + " = 0u;");
            for (std::size_t q = 0; q < parts; ++q) {
                std::string const y = x + "_" + std::to_string(q);
                w_.line(T2() + " dC" + y + " = " + f_.zero2() + ", dS" + y + " = " + f_.zero2() +
                        ";");
            }
        }
        auto body = [&](std::string const &i, std::size_t q) {
            std::string const Q = std::to_string(q);
            w_.line("const " + T2() + " da" + Q + " = " + lds("dj + (" + i + ") * " + u(NR)) +
                    ";");
            w_.line("const " + T2() + " db" + Q + " = " +
                    lds("dj + (" + u(R) + " - (" + i + ")) * " + u(NR)) + ";");
            for (std::size_t b = 0; b < Bu; ++b) {
                std::string const x = std::to_string(b);
                std::string const y = x + "_" + Q;
                w_.line("dm" + x + " += " + kk[b] + ";");
                w_.line("dm" + x + " = dm" + x + " >= " + u(R) + " ? dm" + x + " - " + u(R) +
                        " : dm" + x + ";");
                w_.line("const " + T2() + " dw" + y + " = dts[dm" + x + "];");
                w_.line("dC" + y + " = mk(dC" + y + ".x + da" + Q + ".x * dw" + y + ".x, dC" + y +
                        ".y + da" + Q + ".y * dw" + y + ".x);");
                w_.line("dS" + y + " = mk(dS" + y + ".x + db" + Q + ".x * dw" + y + ".y, dS" + y +
                        ".y + db" + Q + ".y * dw" + y + ".y);");
            }
        };
        std::size_t const full = h / parts * parts; // i = 1..full in steps of `parts`
        if (full > 0) {
            w_.open("for (unsigned di = 1u; di <= " + u(full) + "; di += " + u(parts) + ")");
            for (std::size_t q = 0; q < parts; ++q) {
                w_.open("");
                body(q ? "di + " + u(q) : "di", q);
                w_.close();
            }
            w_.close();
        }
        for (std::size_t i = full + 1; i <= h; ++i) {
            w_.open("");
            body(u(i), 0);
            w_.close();
        }
        // combine: y_k = x_0 + C + dir i S, y_{R-k} = x_0 + C - dir i S
        for (std::size_t b = 0; b < Bu; ++b) {
            std::string const x = std::to_string(b);
            std::string C = "dC" + x + "_0", Sv = "dS" + x + "_0";
            if (parts == 2) {
                w_.line("const " + T2() + " dCt" + x + " = c_add(dC" + x + "_0, dC" + x + "_1);");
                w_.line("const " + T2() + " dSt" + x + " = c_add(dS" + x + "_0, dS" + x + "_1);");
                C = "dCt" + x;
                Sv = "dSt" + x;
            }
            w_.line("const " + T2() + " dc" + x + " = c_add(x0, " + C + ");");
            std::string const c = "dc" + x;
            std::string plus = "mk(" + c + ".x - " + Sv + ".y, " + c + ".y + " + Sv + ".x)";
            std::string minus = "mk(" + c + ".x + " + Sv + ".y, " + c + ".y - " + Sv + ".x)";
            // forward (dir = -1): y_k = C - i S
            w_.line("const " + T2() + " dyk" + x + " = " + (c_.direction < 0 ? minus : plus) +
                    ";");
            w_.line("const " + T2() + " dyr" + x + " = " + (c_.direction < 0 ? plus : minus) +
                    ";");
        }
        // the sums read every slot of the butterfly: all reads end before the outputs land
        if (!dst_global || alias_out_) {
            w_.line(sync());
        }
        auto out_index = [&](std::string const &k) {
            return p == 1 ? "dj * " + u(R) + " + " + k
                          : "(dj / " + u(p) + ") * " + u(p * R) + " + dj % " + u(p) + " + (" + k +
                                ") * " + u(p);
        };
        for (std::size_t b = 0; b < Bu; ++b) {
            std::string const x = std::to_string(b);
            std::string const k = "dk" + x;
            std::string cond = valid[b];
            if (dst_global) {
                cond = cond.empty() ? "active" : "active && " + cond;
            }
            if (cond.empty()) {
                w_.open("");
            } else {
                w_.open("if (" + cond + ")");
            }
            if (dst_global) {
                emit_global_output("dyk" + x, out_index(k));
                w_.open("if (" + k + " != 0u)");
                emit_global_output("dyr" + x, out_index(u(R) + " - " + k));
                w_.close();
            } else {
                w_.line(lds(out_index(k)) + " = dyk" + x + ";");
                w_.line("if (" + k + " != 0u) " + lds(out_index(u(R) + " - " + k)) + " = dyr" + x +
                        ";");
            }
            w_.close();
        }
        w_.close();
        if (!dst_global) {
            w_.line(sync());
        }
    }

    /// c2r: Hermitian input -> complex FFT input z in LDS (reference: sbfft_gen.cpp:221-247
    /// for even N, 318-351 for odd N; the imaginary part of the DC bin is ignored).
    void emit_c2r_pre() {
        if (!input_index_xor_.empty()) {
            throw bad_configuration("permuted input rows are not supported for c2r");
        }
        w_.line("// c2r pre-pass");
        std::size_t const H = nfft_;
        std::size_t const Q = half_ ? H / 2 + 1 : (c_.N - 1) / 2 + 1;
        std::size_t const B = (Q + c_.Nt - 1) / c_.Nt;
        // When the last round holds only the self-conjugate q = H/2 (lane t = 0), its input is
        // 2 conj(X[H/2]): the general formula with the table's twiddle (-1, -0) gives the same
        // values (zeros may differ in sign), at the cost of a full round for one lane.
        bool const lone_mid = half_ && !adjoint() && H % 2 == 0 && (Q - 1) % c_.Nt == 0;
        // phase 1: read the half spectrum and form the complex FFT inputs
        for (std::size_t b = 0; b < B; ++b) {
            std::string q = "q" + std::to_string(b);
            std::string s1 = "s1_" + std::to_string(b), s2 = "s2_" + std::to_string(b);
            w_.line("const unsigned " + q + " = t + " + u(b * c_.Nt) + ";");
            w_.line(T2() + " " + s1 + " = " + f_.zero2() + ", " + s2 + " = " + f_.zero2() + ";");
            if (lone_mid && b + 1 == B) {
                w_.open("if (active && " + q + " == " + u(H / 2) + ")");
                w_.line("const " + T2() + " x = " + gload(ioff(u(H / 2))) + ";");
                w_.line(s1 + " = mk(x.x + x.x, -(x.y + x.y));");
                w_.close();
                continue;
            }
            w_.open("if (active && " + q + " < " + u(Q) + ")");
            if (half_) {
                w_.line(T2() + " x1 = " + gload(ioff(q)) + ";");
                w_.line("if (" + q + " == 0u) x1.y = " + f_.lit(0.0) + ";");
                w_.line(T2() + " x2 = " + gload(ioff(u(H) + " - " + q)) + ";");
                if (adjoint()) {
                    // Hermitian fold: halve the non-self-conjugate bins, drop imag(Nyquist)
                    w_.line("if (" + q + " != 0u) { x1 = c_scale(x1, " + f_.lit(0.5) +
                            "); x2 = c_scale(x2, " + f_.lit(0.5) + "); } else { x2.y = " +
                            f_.lit(0.0) + "; }");
                }
                w_.line("x2.y = -x2.y;");
                w_.line("const " + T2() + " a = c_add(x1, x2);");
                w_.line("const " + T2() + " bt = c_mul(c_sub(x1, x2), " + twr_name() + "[" + q +
                        "]);");
                w_.line(s1 + " = c_add(a, bt);");
                w_.line(s2 + " = mk(a.x - bt.x, bt.y - a.y);");
            } else {
                w_.line(T2() + " ai = " + gload(ioff(q)) + ";");
                w_.line(T2() + " bi = " + f_.zero2() + ";");
                w_.line("if (active_b) bi = " + gload(ioff(q, 1)) + ";");
                w_.line("if (" + q + " == 0u) { ai.y = " + f_.lit(0.0) + "; bi.y = " +
                        f_.lit(0.0) + "; }");
                if (adjoint()) {
                    w_.line("if (" + q + " != 0u) { ai = c_scale(ai, " + f_.lit(0.5) +
                            "); bi = c_scale(bi, " + f_.lit(0.5) + "); }");
                }
                w_.line("const " + T2() + " ib2 = c_ipos(bi);");
                w_.line(s1 + " = c_add(ai, ib2);");
                w_.line("const " + T2() + " d = c_sub(ai, ib2);");
                w_.line(s2 + " = mk(d.x, -d.y);");
            }
            w_.close();
        }
        if (alias_in_) {
            w_.line(sync());
        }
        // phase 2: store them in LDS
        for (std::size_t b = 0; b < B; ++b) {
            std::string q = "q" + std::to_string(b);
            std::string s1 = "s1_" + std::to_string(b), s2 = "s2_" + std::to_string(b);
            w_.open("if (active && " + q + " < " + u(Q) + ")");
            if (half_) {
                w_.line(lds(q + " % " + u(H)) + " = " + s1 + ";");
                w_.line("if (" + q + " != 0u && 2u * " + q + " != " + u(H) + ") " +
                        lds(u(H) + " - " + q) + " = " + s2 + ";");
            } else {
                w_.line(lds(q) + " = " + s1 + ";");
                w_.line("if (" + q + " != 0u) " + lds(u(c_.N) + " - " + q) + " = " + s2 + ";");
            }
            w_.close();
        }
    }

    /// r2c: complex FFT output in LDS -> Hermitian half spectrum (reference: sbfft_gen.cpp:180-199
    /// for even N, 275-291 for odd N).
    void emit_r2c_post() {
        w_.line("// r2c post-pass");
        std::size_t const H = nfft_;
        std::size_t const Q = half_ ? H / 2 + 1 : (c_.N - 1) / 2 + 1;
        std::size_t const B = (Q + c_.Nt - 1) / c_.Nt;
        std::string const half = f_.lit(0.5);
        // When the last round holds only the self-conjugate q = H/2 (lane t = 0), that round is
        // X[H/2] = conj(z[H/2]): the general formula with the table's twiddle (1, 0) gives the
        // same bits, at the cost of a full round for one lane.
        bool const lone_mid = half_ && !adjoint() && H % 2 == 0 && (Q - 1) % c_.Nt == 0;
        // phase 1: read the complex FFT output and form the two spectrum values per q
        for (std::size_t b = 0; b < B; ++b) {
            std::string q = "q" + std::to_string(b);
            std::string o1 = "o1_" + std::to_string(b), o2 = "o2_" + std::to_string(b);
            w_.line("const unsigned " + q + " = t + " + u(b * c_.Nt) + ";");
            w_.line(T2() + " " + o1 + " = " + f_.zero2() + ", " + o2 + " = " + f_.zero2() + ";");
            if (lone_mid && b + 1 == B) {
                w_.open("if (active && " + q + " == " + u(H / 2) + ")");
                w_.line("const " + T2() + " y = " + lds(u(H / 2)) + ";");
                w_.line(o1 + " = mk(y.x, -y.y);");
                w_.close();
                continue;
            }
            w_.open("if (active && " + q + " < " + u(Q) + ")");
            if (half_) {
                w_.line("const " + T2() + " y1 = " + lds(q + " % " + u(H)) + ";");
                w_.line(T2() + " y2 = " + lds("(" + u(H) + " - " + q + ") % " + u(H)) + ";");
                w_.line("y2.y = -y2.y;");
                w_.line("const " + T2() + " a = c_scale(c_add(y2, y1), " + half + ");");
                w_.line("const " + T2() + " bt = c_mul(c_scale(c_sub(y2, y1), " + half + "), " +
                        twr_name() + "[" + q + "]);");
                if (adjoint()) {
                    // doubled non-self-conjugate bins; imag(X[N/2]) = -real(X[0])
                    std::string two = f_.lit(2.0);
                    w_.line("const " + T2() + " z0 = c_add(a, bt);");
                    w_.line("const " + T2() + " z1 = mk(a.x - bt.x, bt.y - a.y);");
                    w_.open("if (" + q + " == 0u)");
                    w_.line(o1 + " = z0;");
                    w_.line(o2 + " = mk(z1.x, -z0.x);");
                    w_.close();
                    w_.open("else");
                    w_.line(o1 + " = c_scale(z0, " + two + ");");
                    w_.line(o2 + " = c_scale(z1, " + two + ");");
                    w_.close();
                } else {
                    w_.line(o1 + " = c_add(a, bt);");
                    w_.line(o2 + " = mk(a.x - bt.x, bt.y - a.y);");
                }
            } else {
                w_.line("const " + T2() + " y1 = " + lds(q) + ";");
                w_.line(T2() + " y2 = " + lds("(" + u(c_.N) + " - " + q + ") % " + u(c_.N)) +
                        ";");
                w_.line("y2.y = -y2.y;");
                // adjoint: doubled non-self-conjugate bins (q >= 1), i.e. no 1/2 there
                std::string w = adjoint() ? "(" + q + " == 0u ? " + half + " : " +
                                                f_.lit(1.0) + ")"
                                          : half;
                w_.line(o1 + " = c_scale(c_add(y2, y1), " + w + ");");
                w_.line("const " + T2() + " d = c_scale(c_sub(y2, y1), " + w + ");");
                w_.line(o2 + " = mk(-d.y, d.x);");
            }
            w_.close();
        }
        if (alias_out_) {
            w_.line(sync());
        }
        // phase 2: store the half spectrum (q = 0 also writes the Nyquist / last bin H)
        for (std::size_t b = 0; b < B; ++b) {
            std::string q = "q" + std::to_string(b);
            std::string o1 = "o1_" + std::to_string(b), o2 = "o2_" + std::to_string(b);
            w_.open("if (active && " + q + " < " + u(Q) + ")");
            if (half_) {
                w_.line(gstore(ooff(q), o1));
                w_.line("if (2u * " + q + " != " + u(H) + ") " + gstore(ooff(u(H) + " - " + q), o2));
            } else {
                w_.line(gstore(ooff(q), o1));
                w_.line("if (active_b) " + gstore(ooff(q, 1), o2));
            }
            w_.close();
        }
    }

    bool adjoint() const { return (c_.flags & fft1d_adjoint) != 0; }

    fft1d_configuration const &c_;
    fp_helper f_;
    std::string name_;
    std::size_t group_size_ = 0;
    bool alias_in_ = false, alias_out_ = false;
    std::string band_;
    writer w_;
    std::size_t nfft_ = 0, L_ = 0, MG_ = 1, np_ = 0;
    unsigned pad_shift_ = 0;
    bool real_ = false, half_ = false, pairs_ = false, use_lds_ = false;
    bool half_in_ = false, bf16_ = false, mfma_ = false;
    bool lds_tables_ = false;   // twiddles read from LDS copies (fused kernels)
    bool lds_barriers_ = false; // barriers order LDS only (fused kernels)
    std::string extra_param_;  // extra parameter of a tile body
    std::string after_input_;  // code run once stage 0 has read all of the tile's input
    std::string input_index_xor_; // XOR-ed into input element indices (permuted input rows)
    bool single_tile_ = false;    // one tile, all columns: `active` is the band alone
    bool lds_table_pad_ = false;  // LDS twiddle rows R + 1 entries apart
    int direct_ = 0; // radix of the direct DFT stage (0: none)
    bool local32_ = false;
    bool prefetch_ = false;
};

} // namespace

/// Complex helpers shared by all kernels of a source; overloaded for float2 and double2.
std::string hip_prelude() {
    std::ostringstream os;
    os << "#ifndef BBFFT_HIP_PRELUDE\n#define BBFFT_HIP_PRELUDE\n";
    for (char const *T : {"float", "double"}) {
        std::string t = T, t2 = t + "2";
        os << "static __device__ __forceinline__ " << t2 << " mk(" << t << " a, " << t
           << " b) { " << t2 << " r; r.x = a; r.y = b; return r; }\n";
        os << "static __device__ __forceinline__ " << t2 << " c_add(" << t2 << " a, " << t2
           << " b) { return mk(a.x + b.x, a.y + b.y); }\n";
        os << "static __device__ __forceinline__ " << t2 << " c_sub(" << t2 << " a, " << t2
           << " b) { return mk(a.x - b.x, a.y - b.y); }\n";
        os << "static __device__ __forceinline__ " << t2 << " c_mul(" << t2 << " a, " << t2
           << " w) { return mk(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x); }\n";
        os << "static __device__ __forceinline__ " << t2 << " c_mulri(" << t2 << " a, " << t
           << " c, " << t << " s) { return mk(a.x * c - a.y * s, a.x * s + a.y * c); }\n";
        os << "static __device__ __forceinline__ " << t2 << " c_scale(" << t2 << " a, " << t
           << " s) { return mk(a.x * s, a.y * s); }\n";
        os << "static __device__ __forceinline__ " << t2 << " c_ipos(" << t2
           << " a) { return mk(-a.y, a.x); }\n";
        os << "static __device__ __forceinline__ " << t2 << " c_ineg(" << t2
           << " a) { return mk(a.y, -a.x); }\n";
        os << "static __device__ __forceinline__ " << t2 << " c_neg(" << t2
           << " a) { return mk(-a.x, -a.y); }\n";
        // streaming (non-temporal) global access, component-wise
        os << "static __device__ __forceinline__ " << t2 << " ld_nt(const " << t2
           << "* p) { return mk(__builtin_nontemporal_load(&p->x), "
              "__builtin_nontemporal_load(&p->y)); }\n";
        os << "static __device__ __forceinline__ void st_nt(" << t2 << "* p, " << t2
           << " v) { __builtin_nontemporal_store(v.x, &p->x); "
              "__builtin_nontemporal_store(v.y, &p->y); }\n";
        os << "static __device__ __forceinline__ " << t << " ld_nt(const " << t
           << "* p) { return __builtin_nontemporal_load(p); }\n";
        os << "static __device__ __forceinline__ void st_nt(" << t << "* p, " << t
           << " v) { __builtin_nontemporal_store(v, p); }\n";
    }
    os << "#endif\n";
    return os.str();
}

namespace {
std::string callback_source(user_module const &um) {
    if (!um) {
        return {};
    }
    std::string src(um.data, um.length);
    if (um.language == kernel_language::opencl_c) {
        return "#define global\n#define __global\n#define kernel\n#define __kernel\n"
               "typedef unsigned int uint;\ntypedef unsigned long ulong;\n"
               "typedef unsigned short ushort;\ntypedef unsigned char uchar;\n"
               "#pragma clang force_cuda_host_device begin\n" +
               src +
               "\n#pragma clang force_cuda_host_device end\n"
               "#undef global\n#undef __global\n#undef kernel\n#undef __kernel\n";
    }
    return src + "\n";
}
} // namespace

void generate_fft1d_body(std::ostream &os, fft1d_configuration const &cfg,
                         std::string const &name, std::size_t group_size) {
    if (cfg.load_function || cfg.store_function) {
        throw bad_configuration("callbacks are not supported in ragged batches");
    }
    if (cfg.type != transform_type::c2c) {
        throw bad_configuration("ragged batches support c2c transforms only");
    }
    kernel_gen g(cfg, name, group_size);
    os << g.generate_body();
}

void generate_fft1d_body_aliased(std::ostream &os, fft1d_configuration const &cfg,
                                 std::string const &name, std::size_t group_size, bool alias_in,
                                 bool alias_out, body_options const &opt) {
    if (cfg.load_function || cfg.store_function) {
        throw bad_configuration("callbacks are not supported in fused kernels");
    }
    if (cfg.flags & (fft1d_input_f16 | fft1d_input_bf16 | fft1d_mfma)) {
        throw bad_configuration("half-precision input is not supported in fused kernels");
    }
    kernel_gen g(cfg, name, group_size, alias_in, alias_out, opt);
    os << g.generate_body();
}

void generate_fft1d(std::ostream &os, fft1d_configuration const &cfg, std::string_view name,
                    user_module const *callbacks) {
    if (cfg.factorization.empty()) {
        throw bad_configuration("empty factorization");
    }
    if ((cfg.load_function || cfg.store_function) && !(callbacks && *callbacks)) {
        throw bad_configuration("callback names given without callback source");
    }
    std::string n = name.empty() ? cfg.identifier() : std::string(name);
    os << hip_prelude();
    if (callbacks) {
        os << callback_source(*callbacks);
    }
    kernel_gen g(cfg, n);
    os << g.generate();
}

} // namespace bbfft
