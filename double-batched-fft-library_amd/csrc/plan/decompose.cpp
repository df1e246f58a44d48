#include "decompose.hpp"
#include "bbfft/bad_configuration.hpp"
#include "bbfft/detail/generator_impl.hpp"

#include <cstdlib>
#include <string>
#include <utility>

namespace bbfft {

namespace {
bool same_strides(std::array<std::size_t, max_tensor_dim> const &a,
                  std::array<std::size_t, max_tensor_dim> const &b, unsigned dim) {
    for (unsigned d = 0; d < dim + 2; ++d) {
        if (a[d] != b[d]) {
            return false;
        }
    }
    return true;
}
} // namespace

pass_layout decompose(configuration const &cfg) {
    if (cfg.dim < 1 || cfg.dim > max_fft_dim) {
        throw bad_configuration("Unsupported FFT dimension: " + std::to_string(cfg.dim));
    }
    if (cfg.istride[0] != 1 || cfg.ostride[0] != 1) {
        throw bad_configuration("Stride s_0 != 1 is not supported");
    }
    if ((cfg.type == transform_type::r2c && cfg.dir != direction::forward) ||
        (cfg.type == transform_type::c2r && cfg.dir != direction::backward)) {
        throw bad_configuration("r2c must be forward and c2r must be backward");
    }
    for (unsigned d = 0; d < cfg.dim + 2; ++d) {
        if (cfg.shape[d] == 0) {
            throw bad_configuration("Tensor shape must not contain zeros");
        }
    }
    pass_layout out;
    if (cfg.dim == 1) {
        configuration c = cfg;
        c.shape = {cfg.shape[0], cfg.shape[1], cfg.shape[2], 0, 0};
        out.passes.push_back(c);
        return out;
    }
    auto is_default = [&](bool inplace) {
        return same_strides(cfg.istride, default_istride(cfg.dim, cfg.shape, cfg.type, inplace),
                            cfg.dim) &&
               same_strides(cfg.ostride, default_ostride(cfg.dim, cfg.shape, cfg.type, inplace),
                            cfg.dim);
    };
    bool const inplace_layout = is_default(true);
    if (!inplace_layout && !is_default(false)) {
        throw bad_configuration("Only default tensor layouts are supported for the nd_fft.");
    }
    bool const real = cfg.type != transform_type::c2c;
    unsigned const D = cfg.dim;
    // Extent of mode d on the complex side and on the real side (in-place r2c pads mode 0).
    auto n_complex = [&](unsigned d) { return d == 0 && real ? cfg.shape[1] / 2 + 1 : cfg.shape[d + 1]; };
    auto n_real = [&](unsigned d) {
        return d == 0 && real && inplace_layout ? 2 * (cfg.shape[1] / 2 + 1) : cfg.shape[d + 1];
    };
    std::size_t prod_complex = 1;
    for (unsigned d = 0; d < D; ++d) {
        prod_complex *= n_complex(d);
    }
    std::vector<configuration> p1d(D);
    std::size_t M = cfg.shape[0];
    std::size_t K = prod_complex * cfg.shape[D + 1];
    for (unsigned d = 0; d < D; ++d) {
        std::size_t const Nd = cfg.shape[d + 1];
        K /= n_complex(d);
        configuration c = {};
        c.dim = 1;
        c.shape = {M, Nd, K, 0, 0};
        c.fp = cfg.fp;
        c.dir = cfg.dir;
        c.type = d == 0 ? cfg.type : transform_type::c2c;
        c.istride = {1, M, M * n_real(d), 0, 0};
        c.ostride = {1, M, M * n_complex(d), 0, 0};
        p1d[d] = c;
        M *= n_complex(d);
    }
    if (cfg.type == transform_type::c2r) {
        for (unsigned d = 0; d < D; ++d) {
            auto c = p1d[D - 1 - d];
            std::swap(c.istride, c.ostride);
            out.passes.push_back(c);
        }
    } else {
        out.passes = std::move(p1d);
    }
    if (cfg.callbacks) {
        // Extension (the reference takes callbacks on 1-D plans only): the load callback runs
        // in the first pass, which reads the plan's input, and the store callback in the last
        // pass, which writes the plan's output, so offsets index the plan's input and output
        // tensors as in a 1-D plan. Passes with callbacks are never fused.
        user_module load = cfg.callbacks, store = cfg.callbacks;
        load.store_function = nullptr;
        store.load_function = nullptr;
        if (load) {
            out.passes.front().callbacks = load;
        }
        if (store) {
            out.passes.back().callbacks = store;
        }
    }
    std::size_t const rb = static_cast<std::size_t>(cfg.fp);
    std::size_t const ib = cfg.type == transform_type::r2c ? rb : 2 * rb;
    std::size_t const ob = cfg.type == transform_type::c2r ? rb : 2 * rb;
    std::size_t const isize = cfg.istride[D + 1] * cfg.shape[D + 1] * ib;
    std::size_t const osize = cfg.ostride[D + 1] * cfg.shape[D + 1] * ob;
    if (isize > osize) {
        out.tmp_bytes = isize;
    }
    return out;
}

bool find_fused_pair(pass_layout const &layout, device_info const &info,
                     unsigned real_pass_variant, fused_pair &out) {
    char const *env = std::getenv("BBFFT_FUSE");
    if (env && std::string(env) == "off") {
        return false;
    }
    bool const forced = env && std::string(env) == "on";
    std::size_t const D = layout.passes.size();
    if (D < 2) {
        return false;
    }
    bool const c2r = layout.passes[D - 1].type == transform_type::c2r;
    std::size_t const first = c2r ? D - 2 : 0;
    auto const &pa = layout.passes[first];
    auto const &pb = layout.passes[first + 1];
    unsigned const va = pa.type == transform_type::c2c ? 0u : real_pass_variant;
    unsigned const vb = pb.type == transform_type::c2c ? 0u : real_pass_variant;
    fused2_configuration fc;
    if (!configure_fused2(pa, pb, info, va, vb, fc)) {
        return false;
    }
    // One work-group per plane. Planes of up to 96 KB of LDS fuse: the 70 KB plane of the
    // 32 x 128^3 fp32 real transforms (one 1024-work-item group per CU) measured 162 us fused
    // against 98 + 91 us for the two passes (r2c) and 169 against 100 + 115 us (c2r) on MI355X.
    // Larger planes (unmeasured) and few planes (under 256: too few groups for the device) are
    // left to the unfused passes, except for problems small enough to be launch-bound (one
    // launch less). BBFFT_FUSE=on fuses whenever it fits.
    std::size_t const total = fc.planes * fc.plane_elems;
    bool const small = total <= (std::size_t(1) << 20);
    if (!forced && !small && (fc.lds_bytes() > 96 * 1024 || fc.planes < 256)) {
        return false;
    }
    // Staging loads with the non-temporal cache policy when every other pass of the plan
    // writes with plain stores: those leave dirty lines in the 256 MB last-level cache, which
    // cache-allocating staging loads would evict (and write back) during the plane kernel.
    // Measured on MI355X (bench.py's plans, processes alternated): 32 x 128^3 fp32 r2c
    // 0.1997 ms against 0.2070, c2r 0.2045 against 0.2142 (third pass with plain stores);
    // plain loads were faster where the other pass streams its stores (c2c 32 x 64 x 128 x 128
    // 0.2574 against 0.2654, r2c 256 x 64^3 0.2293 against 0.2307) and in 2-D plans, whose
    // single kernel then re-reads part of a repeated input from that cache.
    // BBFFT_FUSE_DMA_NT=on|off overrides.
    char const *nt_env = std::getenv("BBFFT_FUSE_DMA_NT");
    if (fc.stage_words && !nt_env) {
        bool plain_stores = D > 2;
        for (std::size_t i = 0; i < D && plain_stores; ++i) {
            if (i == first || i == first + 1) {
                continue;
            }
            try {
                unsigned const v = layout.passes[i].type == transform_type::c2c
                                       ? 0u
                                       : real_pass_variant;
                auto const k = configure_fft1d(layout.passes[i], info, fft1d_tuning{}, v);
                plain_stores = (k.flags & fft1d_nontemporal_store) == 0u;
            } catch (bad_configuration const &) {
                plain_stores = false;
            }
        }
        fc.stage_load_nt = plain_stores;
    }
    out = fused_pair{first, fc};
    return true;
}

configuration adjoint_configuration(configuration const &cfg, unsigned *real_pass_variant) {
    // the configuration's callbacks (if any) apply to the adjoint plan's own input and output
    configuration adj = cfg;
    std::swap(adj.istride, adj.ostride);
    switch (cfg.type) {
    case transform_type::c2c:
        adj.dir = cfg.dir == direction::forward ? direction::backward : direction::forward;
        *real_pass_variant = 0;
        break;
    case transform_type::r2c:
        adj.type = transform_type::c2r;
        adj.dir = direction::backward;
        *real_pass_variant = fft1d_adjoint;
        break;
    case transform_type::c2r:
        adj.type = transform_type::r2c;
        adj.dir = direction::forward;
        *real_pass_variant = fft1d_adjoint;
        break;
    }
    return adj;
}

} // namespace bbfft
