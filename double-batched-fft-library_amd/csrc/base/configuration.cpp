// Default strides and descriptor printing.
// Reference behaviour: src/base/configuration.cpp:24-160.
#include "bbfft/configuration.hpp"

#include <ostream>
#include <sstream>
#include <stdexcept>

namespace bbfft {

char const *to_string(transform_type type) {
    switch (type) {
    case transform_type::c2c:
        return "c2c";
    case transform_type::r2c:
        return "r2c";
    case transform_type::c2r:
        return "c2r";
    }
    return "unknown";
}

auto default_istride(unsigned dim, std::array<std::size_t, max_tensor_dim> const &shape,
                     transform_type type, bool inplace) -> std::array<std::size_t, max_tensor_dim> {
    // Extent of the first FFT mode as laid out in memory.
    std::size_t n1 = shape[1];
    if (type == transform_type::r2c && inplace) {
        n1 = 2 * (shape[1] / 2 + 1);
    } else if (type == transform_type::c2r) {
        n1 = shape[1] / 2 + 1;
    }
    std::array<std::size_t, max_tensor_dim> s{};
    s[0] = 1;
    s[1] = shape[0];
    s[2] = shape[0] * n1;
    for (unsigned d = 2; d <= dim && d + 1 < max_tensor_dim; ++d) {
        s[d + 1] = s[d] * shape[d];
    }
    return s;
}

auto default_ostride(unsigned dim, std::array<std::size_t, max_tensor_dim> const &shape,
                     transform_type type, bool inplace) -> std::array<std::size_t, max_tensor_dim> {
    transform_type mirrored = type;
    if (type == transform_type::r2c) {
        mirrored = transform_type::c2r;
    } else if (type == transform_type::c2r) {
        mirrored = transform_type::r2c;
    }
    return default_istride(dim, shape, mirrored, inplace);
}

void configuration::set_strides_default(bool inplace) {
    istride = default_istride(dim, shape, type, inplace);
    ostride = default_ostride(dim, shape, type, inplace);
}

std::string configuration::to_string() const {
    std::ostringstream oss;
    oss << *this;
    return oss.str();
}

std::ostream &operator<<(std::ostream &os, configuration const &cfg) {
    char prec = 0, domain = 0, dir = 0;
    switch (cfg.fp) {
    case precision::f32:
        prec = 's';
        break;
    case precision::f64:
        prec = 'd';
        break;
    }
    switch (cfg.type) {
    case transform_type::c2c:
        domain = 'c';
        break;
    case transform_type::r2c:
    case transform_type::c2r:
        domain = 'r';
        break;
    }
    switch (cfg.dir) {
    case direction::forward:
        dir = 'f';
        break;
    case direction::backward:
        dir = 'b';
        break;
    }
    if (!prec) {
        throw std::runtime_error("Unsupported precision");
    }
    if (!domain) {
        throw std::runtime_error("Unsupported transform type");
    }
    if (!dir) {
        throw std::runtime_error("Unsupported direction");
    }
    if ((cfg.type == transform_type::r2c && cfg.dir != direction::forward) ||
        (cfg.type == transform_type::c2r && cfg.dir != direction::backward)) {
        throw std::runtime_error(
            "r2c direction must be forward and c2r direction must be backward");
    }

    // In-place means equal strides once both sides are measured in reals for real transforms
    // (the complex side's strides beyond the M mode count complex numbers, i.e. two reals).
    unsigned const nmodes = cfg.dim + 2;
    bool inplace = true;
    for (unsigned d = 0; d < nmodes && d < max_tensor_dim; ++d) {
        std::size_t is = cfg.istride[d], ostr = cfg.ostride[d];
        if (d >= 2 && cfg.type == transform_type::c2r) {
            is *= 2;
        }
        if (d >= 2 && cfg.type == transform_type::r2c) {
            ostr *= 2;
        }
        inplace = inplace && is == ostr;
    }

    os << prec << domain << dir << (inplace ? 'i' : 'o');
    if (cfg.shape[0] != 1) {
        os << cfg.shape[0] << '.';
    }
    os << cfg.shape[1];
    for (unsigned d = 2; d <= cfg.dim; ++d) {
        os << 'x' << cfg.shape[d];
    }
    if (cfg.shape[cfg.dim + 1] != 1) {
        os << '*' << cfg.shape[cfg.dim + 1];
    }
    auto print_stride = [&](char tag, std::array<std::size_t, max_tensor_dim> const &s) {
        os << tag;
        for (unsigned d = 0; d < nmodes; ++d) {
            os << (d ? "," : "") << s[d];
        }
    };
    if (cfg.istride != default_istride(cfg.dim, cfg.shape, cfg.type, inplace)) {
        print_stride('i', cfg.istride);
    }
    if (cfg.ostride != default_ostride(cfg.dim, cfg.shape, cfg.type, inplace)) {
        print_stride('o', cfg.ostride);
    }
    return os;
}

} // namespace bbfft
