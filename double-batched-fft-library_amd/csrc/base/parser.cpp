// FFT descriptor and device-info parsers.
// Reference behaviour: src/base/parser.cpp:59-270 (grammar in docs/manual/descriptor.rst:13-24).
#include "bbfft/parser.hpp"

#include <cctype>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>

namespace bbfft {

namespace {

/// Character cursor that reports errors with a caret under the current position.
class cursor {
  public:
    explicit cursor(std::string_view text) : text_(text) {}

    bool at_end() const { return pos_ >= text_.size(); }
    char peek() const { return at_end() ? '\0' : text_[pos_]; }
    char take() { return at_end() ? '\0' : text_[pos_++]; }
    void skip_blank() {
        while (!at_end() && std::isblank(static_cast<unsigned char>(text_[pos_]))) {
            ++pos_;
        }
    }

    [[noreturn]] void fail(std::string const &msg, bool caret = true) const {
        std::ostringstream oss;
        oss << "==> " << text_ << " is malformed: " << msg;
        if (caret) {
            oss << '\n' << std::string(4 + pos_, ' ') << '^';
        }
        throw std::runtime_error(oss.str());
    }

    std::size_t number() {
        std::size_t start = pos_;
        std::size_t v = 0;
        // numbers are ints, as the reference reads them (std::stoi, src/base/parser.cpp:48)
        std::size_t const max = static_cast<std::size_t>(std::numeric_limits<int>::max());
        while (!at_end() && std::isdigit(static_cast<unsigned char>(text_[pos_]))) {
            std::size_t const d = static_cast<std::size_t>(text_[pos_] - '0');
            if (v > (max - d) / 10) {
                fail("number out of range");
            }
            v = 10 * v + d;
            ++pos_;
        }
        if (start == pos_) {
            fail("expected number");
        }
        return v;
    }

    /// Consume one of the characters in `set`, or fail with `what`.
    char one_of(char const *set, char const *what) {
        char c = peek();
        for (char const *s = set; *s; ++s) {
            if (c == *s && c != '\0') {
                ++pos_;
                return c;
            }
        }
        fail(std::string("expected ") + what);
    }

  private:
    std::string_view text_;
    std::size_t pos_ = 0;
};

} // namespace

configuration parse_fft_descriptor(std::string_view desc) {
    cursor cur(desc);
    configuration cfg = {};

    cfg.fp = cur.one_of("sd", "'s' (single) or 'd' (double)") == 's' ? precision::f32
                                                                      : precision::f64;
    bool const real = cur.one_of("cr", "'c' (complex) or 'r' (real)") == 'r';
    bool const forward = cur.one_of("fb", "'f' (forward) or 'b' (backward)") == 'f';
    bool const inplace = cur.one_of("io", "'i' (in-place) or 'o' (out-of-place)") == 'i';
    cfg.dir = forward ? direction::forward : direction::backward;
    cfg.type = real ? (forward ? transform_type::r2c : transform_type::c2r) : transform_type::c2c;

    // Shape: numbers separated by '.', 'x' or '*'; '.' may only follow the first number (M)
    // and '*' may only precede the last one (K).
    std::size_t numbers[max_tensor_dim] = {};
    char seps[max_tensor_dim] = {};
    unsigned count = 0;
    numbers[count++] = cur.number();
    while (!cur.at_end()) {
        char c = cur.peek();
        if (c != '.' && c != 'x' && c != '*') {
            break;
        }
        if (count >= max_tensor_dim) {
            cur.fail("tensor dimension must not be larger than " + std::to_string(max_tensor_dim),
                     false);
        }
        cur.take();
        seps[count - 1] = c;
        numbers[count++] = cur.number();
    }
    unsigned const nsep = count - 1;
    bool const has_M = nsep >= 1 && seps[0] == '.';
    bool const has_K = nsep >= 1 && seps[nsep - 1] == '*';
    unsigned fft_dim = 1;
    for (unsigned i = has_M ? 1 : 0; i + (has_K ? 1 : 0) < nsep; ++i) {
        if (seps[i] != 'x') {
            cur.fail("'.' or '*' must only appear at the beginning or end of the tensor shape, "
                     "respectively",
                     false);
        }
        ++fft_dim;
    }
    if (fft_dim > max_fft_dim) {
        cur.fail("only " + std::to_string(max_fft_dim - 1) + " 'x' are supported", false);
    }
    cfg.dim = fft_dim;
    unsigned idx = 0;
    cfg.shape = {};
    cfg.shape[0] = has_M ? numbers[idx++] : 1;
    for (unsigned d = 1; d <= fft_dim; ++d) {
        cfg.shape[d] = numbers[idx++];
    }
    cfg.shape[fft_dim + 1] = has_K ? numbers[idx++] : 1;

    auto stride_list = [&](std::array<std::size_t, max_tensor_dim> &s) {
        s = {};
        for (unsigned d = 0; d < fft_dim + 2; ++d) {
            if (d > 0) {
                if (cur.peek() != ',') {
                    cur.fail("expected ,");
                }
                cur.take();
            }
            s[d] = cur.number();
        }
    };
    bool custom_i = false, custom_o = false;
    while (!cur.at_end()) {
        char c = cur.one_of("io", "'i' (istride) or 'o' (ostride)");
        if (c == 'i') {
            stride_list(cfg.istride);
            custom_i = true;
        } else {
            stride_list(cfg.ostride);
            custom_o = true;
        }
    }
    if (!custom_i) {
        cfg.istride = default_istride(cfg.dim, cfg.shape, cfg.type, inplace);
    }
    if (!custom_o) {
        cfg.ostride = default_ostride(cfg.dim, cfg.shape, cfg.type, inplace);
    }
    return cfg;
}

device_info parse_device_info(std::string_view desc) {
    cursor cur(desc);
    device_info info = {};
    auto expect = [&](char c) {
        cur.skip_blank();
        if (cur.peek() != c) {
            cur.fail(std::string("expected '") + c + "'");
        }
        cur.take();
    };
    auto num = [&]() {
        cur.skip_blank();
        return cur.number();
    };
    expect('{');
    info.max_work_group_size = num();
    expect(',');
    expect('{');
    info.subgroup_sizes.push_back(num());
    for (;;) {
        cur.skip_blank();
        if (cur.peek() == ',') {
            cur.take();
            info.subgroup_sizes.push_back(num());
        } else {
            break;
        }
    }
    expect('}');
    expect(',');
    info.local_memory_size = num();
    expect(',');
    cur.skip_blank();
    if (cur.peek() == 'g') {
        cur.take();
        expect('p');
        expect('u');
        info.type = device_type::gpu;
    } else if (cur.peek() == 'c') {
        cur.take();
        expect('p');
        expect('u');
        info.type = device_type::cpu;
    } else {
        cur.fail("expected gpu or cpu");
    }
    expect('}');
    return info;
}

} // namespace bbfft
