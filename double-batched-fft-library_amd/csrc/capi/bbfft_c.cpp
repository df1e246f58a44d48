// C ABI over the C++ API (declarations and reference citations in include/bbfft_hip.h).
#include "bbfft_hip.h"

#include "../base/math.hpp"
#include "../generator/roots.hpp"
#include "../hip/aot_archive.hpp"
#include "../hip/hip_plan_impl.hpp"
#include "../hip/static_scratch.hpp"
#include "../plan/decompose.hpp"
#include "bbfft/aot_cache.hpp"
#include "bbfft/bad_configuration.hpp"
#include "bbfft/configuration.hpp"
#include "bbfft/detail/generator_impl.hpp"
#include "bbfft/hip/device.hpp"
#include "bbfft/hip/error.hpp"
#include "bbfft/hip/make_plan.hpp"
#include "bbfft/hip/online_compiler.hpp"
#include "bbfft/jit_cache_all.hpp"
#include "bbfft/parser.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <sstream>
#include <string>
#include <vector>

using namespace bbfft;

struct bbfft_plan_s {
    std::shared_ptr<hip_detail::plan_base> impl;
};

/// AOT modules first, then modules compiled in this process.
struct bbfft_cache_s : public jit_cache {
    aot_cache aot;
    jit_cache_all jit;
    std::size_t aot_kernels = 0;
    auto get(jit_cache_key const &key) const -> shared_handle<module_handle_t> override {
        auto m = aot.get(key);
        return m ? m : jit.get(key);
    }
    void store(jit_cache_key const &key, shared_handle<module_handle_t> mod) override {
        jit.store(key, std::move(mod));
    }
};

namespace {

thread_local std::string g_last_error;

template <typename F> int guarded(F &&f) {
    try {
        g_last_error.clear();
        f();
        return BBFFT_SUCCESS;
    } catch (bad_configuration const &e) {
        g_last_error = e.what();
        return BBFFT_ERROR_BAD_CONFIGURATION;
    } catch (hip::error const &e) {
        g_last_error = e.what();
        return BBFFT_ERROR_RUNTIME;
    } catch (std::invalid_argument const &e) {
        g_last_error = e.what();
        return BBFFT_ERROR_INVALID_ARGUMENT;
    } catch (std::runtime_error const &e) {
        g_last_error = e.what();
        return BBFFT_ERROR_PARSE;
    } catch (std::exception const &e) {
        g_last_error = e.what();
        return BBFFT_ERROR_INTERNAL;
    } catch (...) {
        g_last_error = "unknown error";
        return BBFFT_ERROR_INTERNAL;
    }
}

void require(bool ok, char const *what) {
    if (!ok) {
        throw std::invalid_argument(what);
    }
}

configuration from_c(bbfft_configuration const &c) {
    configuration cfg = {};
    cfg.dim = c.dim;
    for (unsigned i = 0; i < max_tensor_dim; ++i) {
        cfg.shape[i] = c.shape[i];
        cfg.istride[i] = c.istride[i];
        cfg.ostride[i] = c.ostride[i];
    }
    require(c.precision == 4 || c.precision == 8, "precision must be 4 or 8");
    require(c.direction == -1 || c.direction == 1, "direction must be -1 or +1");
    require(c.type >= 0 && c.type <= 2, "type must be 0 (c2c), 1 (r2c) or 2 (c2r)");
    cfg.fp = static_cast<precision>(c.precision);
    cfg.dir = static_cast<direction>(c.direction);
    cfg.type = static_cast<transform_type>(c.type);
    cfg.callbacks.data = c.callback_source;
    cfg.callbacks.length = c.callback_length;
    cfg.callbacks.load_function = c.load_function;
    cfg.callbacks.store_function = c.store_function;
    cfg.callbacks.language =
        c.callback_language == 1 ? kernel_language::hip : kernel_language::opencl_c;
    return cfg;
}

void to_c(configuration const &cfg, bbfft_configuration &c) {
    std::memset(&c, 0, sizeof(c));
    c.dim = cfg.dim;
    for (unsigned i = 0; i < max_tensor_dim; ++i) {
        c.shape[i] = cfg.shape[i];
        c.istride[i] = cfg.istride[i];
        c.ostride[i] = cfg.ostride[i];
    }
    c.precision = static_cast<int>(cfg.fp);
    c.direction = static_cast<int>(cfg.dir);
    c.type = static_cast<int>(cfg.type);
}

void copy_out(std::string const &s, char *buf, std::size_t len) {
    require(buf != nullptr || len == 0, "null output buffer");
    if (s.size() + 1 > len) {
        throw std::invalid_argument("output buffer too small (" + std::to_string(s.size() + 1) +
                                    " bytes needed)");
    }
    std::memcpy(buf, s.c_str(), s.size() + 1);
}

std::array<std::size_t, max_tensor_dim> shape_of(size_t const *shape) {
    std::array<std::size_t, max_tensor_dim> s{};
    for (unsigned i = 0; i < max_tensor_dim; ++i) {
        s[i] = shape[i];
    }
    return s;
}

fft1d_tuning tuning_of(bbfft_tuning const *t) {
    fft1d_tuning r;
    if (!t) {
        return r;
    }
    require(t->num_factors >= 0 && t->num_factors <= 8, "num_factors must be in [0, 8]");
    r.factorization.assign(t->factorization, t->factorization + t->num_factors);
    r.Nt = t->Nt;
    r.Mb = t->Mb;
    r.Kb = t->Kb;
    r.lds_pad = t->lds_pad;
    r.no_pad = t->no_pad != 0;
    r.flags = t->flags;
    r.has_flags = t->has_flags != 0;
    r.waves_per_eu = t->waves_per_eu;
    r.groups_per_cu = t->groups_per_cu;
    return r;
}

fft1d_configuration kernel_config_of(bbfft_configuration const *cfg, unsigned pass,
                                     char const *device_info, configuration *pass_cfg,
                                     bbfft_tuning const *tuning = nullptr) {
    require(cfg && device_info, "null argument");
    auto c = from_c(*cfg);
    auto layout = decompose(c);
    require(pass < layout.passes.size(), "pass index out of range");
    if (pass_cfg) {
        *pass_cfg = layout.passes[pass];
        pass_cfg->callbacks = c.callbacks;
    }
    require(tuning == nullptr || layout.passes.size() == 1, "tuning applies to 1-D only");
    return configure_fft1d(layout.passes[pass], parse_device_info(device_info),
                           tuning_of(tuning));
}

ragged_configuration ragged_of(int prec, bbfft_ragged_job const *jobs, size_t num_jobs) {
    require(jobs != nullptr || num_jobs == 0, "null jobs");
    require(prec == 4 || prec == 8, "precision must be 4 or 8");
    ragged_configuration r;
    r.fp = static_cast<precision>(prec);
    r.jobs.resize(num_jobs);
    for (size_t j = 0; j < num_jobs; ++j) {
        require(jobs[j].direction == -1 || jobs[j].direction == 1,
                "ragged job direction must be -1 or +1");
        r.jobs[j] = ragged_job{jobs[j].M, jobs[j].N, jobs[j].K,
                               static_cast<direction>(jobs[j].direction), jobs[j].in_offset,
                               jobs[j].out_offset};
    }
    return r;
}

} // namespace

extern "C" {

const char *bbfft_last_error(void) { return g_last_error.c_str(); }

const char *bbfft_version(void) { return "0.5.1-gfx950"; }

int bbfft_parse_fft_descriptor(const char *desc, bbfft_configuration *cfg) {
    return guarded([&] {
        require(desc && cfg, "null argument");
        to_c(parse_fft_descriptor(desc), *cfg);
    });
}

int bbfft_configuration_to_string(const bbfft_configuration *cfg, char *buf, size_t len) {
    return guarded([&] {
        require(cfg != nullptr, "null configuration");
        copy_out(from_c(*cfg).to_string(), buf, len);
    });
}

int bbfft_default_istride(unsigned dim, const size_t shape[5], int type, int inplace,
                          size_t stride[5]) {
    return guarded([&] {
        require(shape && stride && type >= 0 && type <= 2, "invalid argument");
        auto s = default_istride(dim, shape_of(shape), static_cast<transform_type>(type), inplace);
        std::copy(s.begin(), s.end(), stride);
    });
}

int bbfft_default_ostride(unsigned dim, const size_t shape[5], int type, int inplace,
                          size_t stride[5]) {
    return guarded([&] {
        require(shape && stride && type >= 0 && type <= 2, "invalid argument");
        auto s = default_ostride(dim, shape_of(shape), static_cast<transform_type>(type), inplace);
        std::copy(s.begin(), s.end(), stride);
    });
}

int bbfft_device_info_normalize(const char *desc, char *buf, size_t len) {
    return guarded([&] {
        require(desc != nullptr, "null descriptor");
        copy_out(parse_device_info(desc).to_string(), buf, len);
    });
}

int bbfft_factor(unsigned n, unsigned index, unsigned *factors) {
    return guarded([&] {
        require(factors != nullptr || index == 0, "null output");
        auto f = factor(n, index);
        std::copy(f.begin(), f.end(), factors);
    });
}

unsigned bbfft_iroot(unsigned radicand, unsigned index) { return iroot(radicand, index); }

int bbfft_is_prime(unsigned n) { return is_prime(n) ? 1 : 0; }

int bbfft_root_of_unity(long k, long n, int precision, double *re, double *im) {
    return guarded([&] {
        require(re && im && n > 0 && (precision == 4 || precision == 8), "invalid argument");
        if (precision == 4) {
            auto w = gen::root_of_unity_f(k, n);
            *re = w.real();
            *im = w.imag();
        } else {
            auto w = gen::root_of_unity(k, n);
            *re = w.real();
            *im = w.imag();
        }
    });
}

int bbfft_num_passes(const bbfft_configuration *cfg, unsigned *num) {
    return guarded([&] {
        require(cfg && num, "null argument");
        *num = static_cast<unsigned>(decompose(from_c(*cfg)).passes.size());
    });
}

namespace {
void fill_kernel_config(fft1d_configuration const &kc, configuration const &pc,
                        bbfft_kernel_config *out, char *name, size_t name_len);
}

int bbfft_kernel_configuration_tuned(const bbfft_configuration *cfg, const bbfft_tuning *tuning,
                                     const char *device_info, bbfft_kernel_config *out,
                                     char *name, size_t name_len) {
    return guarded([&] {
        require(out != nullptr, "null output");
        configuration pc = {};
        auto kc = kernel_config_of(cfg, 0, device_info, &pc, tuning);
        fill_kernel_config(kc, pc, out, name, name_len);
    });
}

int bbfft_generate_source_tuned(const bbfft_configuration *cfg, const bbfft_tuning *tuning,
                                const char *device_info, char *buf, size_t len, size_t *needed) {
    return guarded([&] {
        configuration pc = {};
        auto kc = kernel_config_of(cfg, 0, device_info, &pc, tuning);
        std::ostringstream os;
        generate_fft1d(os, kc, {}, &pc.callbacks);
        auto s = os.str();
        if (needed) {
            *needed = s.size() + 1;
        }
        if (buf) {
            copy_out(s, buf, len);
        }
    });
}

int bbfft_adjoint_configuration(const bbfft_configuration *cfg, bbfft_configuration *adjoint) {
    return guarded([&] {
        require(cfg && adjoint, "null argument");
        unsigned variant = 0;
        to_c(adjoint_configuration(from_c(*cfg), &variant), *adjoint);
    });
}

int bbfft_plan_create_adjoint(const bbfft_configuration *cfg, bbfft_stream_t stream,
                              bbfft_cache_t cache, bbfft_plan_t *plan) {
    return guarded([&] {
        require(cfg && plan, "null argument");
        auto impl = hip_detail::create_adjoint(from_c(*cfg), static_cast<hipStream_t>(stream),
                                               cache);
        *plan = new bbfft_plan_s{std::move(impl)};
    });
}

int bbfft_plan_create_tuned(const bbfft_configuration *cfg, const bbfft_tuning *tuning,
                            bbfft_stream_t stream, bbfft_cache_t cache, bbfft_plan_t *plan) {
    return guarded([&] {
        require(cfg && tuning && plan, "null argument");
        auto impl = hip_detail::create_tuned(from_c(*cfg), tuning_of(tuning),
                                             static_cast<hipStream_t>(stream), cache);
        *plan = new bbfft_plan_s{std::move(impl)};
    });
}

int bbfft_kernel_configuration(const bbfft_configuration *cfg, unsigned pass,
                               const char *device_info, bbfft_kernel_config *out, char *name,
                               size_t name_len) {
    return guarded([&] {
        require(out != nullptr, "null output");
        configuration pc = {};
        auto kc = kernel_config_of(cfg, pass, device_info, &pc);
        fill_kernel_config(kc, pc, out, name, name_len);
    });
}

namespace {
void fill_kernel_config(fft1d_configuration const &kc, configuration const &pc,
                        bbfft_kernel_config *out, char *name, size_t name_len) {
    std::memset(out, 0, sizeof(*out));
    out->M = kc.M;
    out->K = pc.shape[2];
    out->Mb = kc.Mb;
    out->N = kc.N;
    out->Nt = kc.Nt;
    out->Kb = kc.Kb;
    out->lds_pad = kc.lds_pad;
    out->lds_bytes = kc.lds_bytes();
    out->work_group_size = kc.work_group_size();
    out->num_factors = static_cast<int>(std::min<std::size_t>(8, kc.factorization.size()));
    for (int i = 0; i < out->num_factors; ++i) {
        out->factorization[i] = kc.factorization[i];
    }
    out->inplace_unsupported = kc.inplace_unsupported ? 1 : 0;
    if (name) {
        copy_out(kc.identifier(), name, name_len);
    }
}
} // namespace

int bbfft_generate_source(const bbfft_configuration *cfg, unsigned pass, const char *device_info,
                          char *buf, size_t len, size_t *needed) {
    return guarded([&] {
        configuration pc = {};
        auto kc = kernel_config_of(cfg, pass, device_info, &pc);
        std::ostringstream os;
        generate_fft1d(os, kc, {}, &pc.callbacks);
        auto s = os.str();
        if (needed) {
            *needed = s.size() + 1;
        }
        if (buf) {
            copy_out(s, buf, len);
        }
    });
}

int bbfft_compile_kernel(const bbfft_configuration *cfg, unsigned pass, const char *device_info,
                         const char *arch, size_t *code_size) {
    return guarded([&] {
        require(arch != nullptr, "null arch");
        configuration pc = {};
        auto kc = kernel_config_of(cfg, pass, device_info, &pc);
        std::ostringstream os;
        generate_fft1d(os, kc, {}, &pc.callbacks);
        auto code = hip::compile_to_native(os.str(), arch);
        if (code_size) {
            *code_size = code.size();
        }
    });
}

int bbfft_cache_create(bbfft_cache_t *cache) {
    return guarded([&] {
        require(cache != nullptr, "null output");
        *cache = new bbfft_cache_s();
    });
}

int bbfft_cache_load_aot(bbfft_cache_t cache, const char *path, int device, size_t *num) {
    return guarded([&] {
        require(cache && path, "null argument");
        auto n = hip::register_aot_archive(cache->aot, path, device);
        cache->aot_kernels += n;
        if (num) {
            *num = n;
        }
    });
}

int bbfft_cache_load_aot_memory(bbfft_cache_t cache, const void *data, size_t size, int device,
                                size_t *num) {
    return guarded([&] {
        require(cache && data, "null argument");
        auto n = hip::register_aot_archive(cache->aot, static_cast<std::uint8_t const *>(data),
                                           size, device);
        cache->aot_kernels += n;
        if (num) {
            *num = n;
        }
    });
}

int bbfft_aot_archive_kernels(const void *data, size_t size, char *buf, size_t len,
                              size_t *needed) {
    return guarded([&] {
        require(data != nullptr, "null argument");
        auto entries =
            hip::parse_archive(static_cast<std::uint8_t const *>(data), size, "(memory)");
        std::string s;
        for (auto const &e : entries) {
            for (auto const &n : e.kernel_names) {
                s += e.arch + " " + n + "\n";
            }
        }
        if (needed) {
            *needed = s.size() + 1;
        }
        if (buf) {
            copy_out(s, buf, len);
        }
    });
}

int bbfft_cache_size(bbfft_cache_t cache, size_t *num) {
    return guarded([&] {
        require(cache && num, "null argument");
        *num = cache->jit.kernel_names().size() + cache->aot_kernels;
    });
}

void bbfft_cache_destroy(bbfft_cache_t cache) { delete cache; }

int bbfft_plan_create(const bbfft_configuration *cfg, bbfft_stream_t stream, bbfft_cache_t cache,
                      bbfft_plan_t *plan) {
    return guarded([&] {
        require(cfg && plan, "null argument");
        auto impl = hip_detail::create(from_c(*cfg), static_cast<hipStream_t>(stream), cache);
        *plan = new bbfft_plan_s{std::move(impl)};
    });
}

int bbfft_plan_create_from_descriptor(const char *desc, bbfft_stream_t stream,
                                      bbfft_cache_t cache, bbfft_plan_t *plan) {
    return guarded([&] {
        require(desc && plan, "null argument");
        auto impl =
            hip_detail::create(parse_fft_descriptor(desc), static_cast<hipStream_t>(stream), cache);
        *plan = new bbfft_plan_s{std::move(impl)};
    });
}

int bbfft_plan_execute(bbfft_plan_t plan, const void *in, void *out, bbfft_event_t signal,
                       uint32_t num_wait, const bbfft_event_t *wait) {
    return guarded([&] {
        require(plan != nullptr, "null plan");
        require(num_wait == 0 || wait != nullptr, "wait events missing");
        std::vector<hipEvent_t> w(num_wait);
        for (uint32_t i = 0; i < num_wait; ++i) {
            w[i] = static_cast<hipEvent_t>(wait[i]);
        }
        plan->impl->execute(in, out, static_cast<hipEvent_t>(signal), num_wait,
                            num_wait ? w.data() : nullptr);
    });
}

int bbfft_plan_execute_static_scratch(bbfft_plan_t plan, const void *in, void *out,
                                      bbfft_event_t signal, uint32_t num_wait,
                                      const bbfft_event_t *wait) {
    return guarded([&] {
        require(plan != nullptr, "null plan");
        require(num_wait == 0 || wait != nullptr, "wait events missing");
        std::size_t const need = plan->impl->scratch_bytes();
        if (need > hip_detail::static_scratch_bytes) {
            throw bad_configuration(
                "The plan needs " + std::to_string(need) +
                " bytes of scratch memory, more than the static scratch array holds (" +
                std::to_string(hip_detail::static_scratch_bytes) +
                " bytes); use bbfft_plan_execute");
        }
        void *scratch = nullptr;
        if (need) {
            scratch = hip_detail::static_scratch(plan->impl->device());
        }
        hipStream_t const stream = plan->impl->stream();
        for (uint32_t i = 0; i < num_wait; ++i) {
            BBFFT_HIP_CHECK(hipStreamWaitEvent(stream, static_cast<hipEvent_t>(wait[i]), 0));
        }
        plan->impl->execute_with_scratch(in, out, scratch);
        if (signal) {
            BBFFT_HIP_CHECK(hipEventRecord(static_cast<hipEvent_t>(signal), stream));
        }
    });
}

int bbfft_plan_num_launches(bbfft_plan_t plan, unsigned *num) {
    return guarded([&] {
        require(plan && num, "null argument");
        *num = static_cast<unsigned>(plan->impl->launches().size());
    });
}

int bbfft_plan_launch_info(bbfft_plan_t plan, unsigned index, char *name, size_t name_len,
                           unsigned *grid, unsigned *block, size_t *lds_bytes) {
    return guarded([&] {
        require(plan != nullptr, "null plan");
        auto l = plan->impl->launches();
        require(index < l.size(), "launch index out of range");
        if (name) {
            copy_out(l[index].kernel_name, name, name_len);
        }
        if (grid) {
            *grid = l[index].grid;
        }
        if (block) {
            *block = l[index].block;
        }
        if (lds_bytes) {
            *lds_bytes = l[index].lds_bytes;
        }
    });
}

void bbfft_plan_destroy(bbfft_plan_t plan) { delete plan; }

int bbfft_device_info_query(int device, char *buf, size_t len) {
    return guarded([&] { copy_out(hip::get_device_info(device).to_string(), buf, len); });
}

int bbfft_fused_configuration(const bbfft_configuration *cfg, const char *device_info,
                              int *fused, bbfft_fused_info *info, char *name, size_t name_len) {
    return guarded([&] {
        require(cfg && device_info && fused, "null argument");
        auto layout = decompose(from_c(*cfg));
        fused_pair fp;
        *fused = find_fused_pair(layout, parse_device_info(device_info), 0u, fp) ? 1 : 0;
        if (*fused && info) {
            info->first_pass = static_cast<unsigned>(fp.first);
            info->group_size = fp.fc.group_size;
            info->lds_bytes = fp.fc.lds_bytes() + 4 * fp.fc.stage_words;
            info->planes = fp.fc.planes;
            info->ka = fp.fc.ka;
            info->kb = fp.fc.kb;
            info->plane_elems = fp.fc.plane_elems;
            info->in_plane_stride = fp.fc.in_plane_stride;
            info->out_plane_stride = fp.fc.out_plane_stride;
        }
        if (*fused && name) {
            copy_out(fp.fc.identifier(), name, name_len);
        }
    });
}

int bbfft_generate_fused_source(const bbfft_configuration *cfg, const char *device_info,
                                char *buf, size_t len, size_t *needed) {
    return guarded([&] {
        require(cfg && device_info, "null argument");
        auto layout = decompose(from_c(*cfg));
        fused_pair fp;
        require(find_fused_pair(layout, parse_device_info(device_info), 0u, fp),
                "the configuration has no fused passes on this device");
        std::ostringstream os;
        generate_fused2(os, fp.fc);
        auto str = os.str();
        if (needed) {
            *needed = str.size() + 1;
        }
        if (buf) {
            copy_out(str, buf, len);
        }
    });
}

int bbfft_half_input_kernel_configuration(const bbfft_configuration *cfg, unsigned flags,
                                          const char *device_info, bbfft_kernel_config *out,
                                          char *name, size_t name_len) {
    return guarded([&] {
        require(cfg && device_info && out, "null argument");
        unsigned const allowed = BBFFT_INPUT_F16 | BBFFT_INPUT_BF16 | BBFFT_MFMA;
        require((flags & ~allowed) == 0 && (flags & (BBFFT_INPUT_F16 | BBFFT_INPUT_BF16)),
                "invalid half-input flags");
        auto layout = decompose(from_c(*cfg));
        require(layout.passes.size() == 1, "half-input plans are 1-D");
        auto kc = configure_fft1d(layout.passes[0], parse_device_info(device_info), {}, flags);
        fill_kernel_config(kc, layout.passes[0], out, name, name_len);
    });
}

int bbfft_plan_create_half_input(const bbfft_configuration *cfg, unsigned flags,
                                 bbfft_stream_t stream, bbfft_cache_t cache, bbfft_plan_t *plan) {
    return guarded([&] {
        require(cfg && plan, "null argument");
        unsigned const allowed = BBFFT_INPUT_F16 | BBFFT_INPUT_BF16 | BBFFT_MFMA;
        require((flags & ~allowed) == 0, "unknown half-input flags");
        require((flags & (BBFFT_INPUT_F16 | BBFFT_INPUT_BF16)) != 0,
                "one of BBFFT_INPUT_F16 / BBFFT_INPUT_BF16 is required");
        auto impl = hip_detail::create_variant(from_c(*cfg), flags,
                                               static_cast<hipStream_t>(stream), cache);
        *plan = new bbfft_plan_s{std::move(impl)};
    });
}

int bbfft_ragged_kernel_configuration(int prec, const bbfft_ragged_job *jobs, size_t num_jobs,
                                      const char *device_info, bbfft_ragged_info *info,
                                      char *name, size_t name_len) {
    return guarded([&] {
        require(device_info && info, "null argument");
        auto rk = configure_ragged(ragged_of(prec, jobs, num_jobs), parse_device_info(device_info));
        info->group_size = rk.group_size;
        info->lds_bytes = rk.lds_bytes();
        info->num_variants = rk.variants.size();
        info->total_tiles = rk.total_tiles();
        if (name) {
            copy_out(rk.identifier(), name, name_len);
        }
    });
}

int bbfft_ragged_variant_configuration(int prec, const bbfft_ragged_job *jobs, size_t num_jobs,
                                       const char *device_info, size_t variant,
                                       bbfft_kernel_config *out) {
    return guarded([&] {
        require(device_info && out, "null argument");
        auto rk = configure_ragged(ragged_of(prec, jobs, num_jobs), parse_device_info(device_info));
        require(variant < rk.variants.size(), "variant index out of range");
        auto const &kc = rk.variants[variant];
        configuration pc{1, {kc.M, kc.N, 0}, kc.fp, static_cast<direction>(kc.direction),
                         transform_type::c2c};
        fill_kernel_config(kc, pc, out, nullptr, 0);
    });
}

int bbfft_ragged_job_variants(int prec, const bbfft_ragged_job *jobs, size_t num_jobs,
                              const char *device_info, unsigned *job_variant) {
    return guarded([&] {
        require(device_info && job_variant, "null argument");
        auto rk = configure_ragged(ragged_of(prec, jobs, num_jobs), parse_device_info(device_info));
        std::copy(rk.job_variant.begin(), rk.job_variant.end(), job_variant);
    });
}

int bbfft_ragged_job_table(int prec, const bbfft_ragged_job *jobs, size_t num_jobs,
                           const char *device_info, size_t grid, unsigned long long *table,
                           size_t len, size_t *needed) {
    return guarded([&] {
        require(device_info != nullptr, "null argument");
        require(grid > 0, "grid must be positive");
        auto r = ragged_of(prec, jobs, num_jobs);
        auto t = ragged_job_table(r, configure_ragged(r, parse_device_info(device_info)), grid);
        if (needed) {
            *needed = t.size();
        }
        if (table) {
            require(len >= t.size(), "table buffer too small");
            std::copy(t.begin(), t.end(), table);
        }
    });
}

int bbfft_generate_ragged_source(int prec, const bbfft_ragged_job *jobs, size_t num_jobs,
                                 const char *device_info, char *buf, size_t len,
                                 size_t *needed) {
    return guarded([&] {
        require(device_info != nullptr, "null argument");
        auto rk = configure_ragged(ragged_of(prec, jobs, num_jobs), parse_device_info(device_info));
        std::ostringstream os;
        generate_ragged(os, rk);
        auto s = os.str();
        if (needed) {
            *needed = s.size() + 1;
        }
        if (buf) {
            copy_out(s, buf, len);
        }
    });
}

int bbfft_ragged_plan_create(int prec, const bbfft_ragged_job *jobs, size_t num_jobs,
                             unsigned flags, bbfft_stream_t stream, bbfft_cache_t cache,
                             bbfft_plan_t *plan) {
    return guarded([&] {
        require(plan != nullptr, "null argument");
        require((flags & ~unsigned(BBFFT_RAGGED_GRAPH)) == 0, "unknown ragged plan flags");
        auto r = ragged_of(prec, jobs, num_jobs);
        int device = 0;
        BBFFT_HIP_CHECK(hipGetDevice(&device));
        auto impl = std::make_shared<hip_detail::ragged_plan>(
            r, static_cast<hipStream_t>(stream), device, cache,
            (flags & BBFFT_RAGGED_GRAPH) != 0);
        *plan = new bbfft_plan_s{std::move(impl)};
    });
}

int bbfft_aot_generate_ragged(int prec, const bbfft_ragged_job *jobs, size_t num_jobs,
                              const char *device_info, const char *arch, const char *out_path) {
    return guarded([&] {
        require(device_info && arch && out_path, "null argument");
        auto rk = configure_ragged(ragged_of(prec, jobs, num_jobs), parse_device_info(device_info));
        std::ostringstream os;
        generate_ragged(os, rk);
        std::vector<hip::archive_entry> entries(2);
        entries[0].arch = arch;
        entries[0].kernel_names.push_back(rk.identifier());
        entries[0].code = hip::compile_to_native(os.str(), arch);
        // The plain kernel of each variant with a K tile of 1 and no access hints (the same
        // arithmetic per job).
        std::ostringstream ps;
        std::set<std::string> seen;
        for (auto const &v : rk.variants) {
            auto p = v;
            p.Kb = 1;
            p.flags = 0;
            if (seen.insert(p.identifier()).second) {
                generate_fft1d(ps, p);
                entries[1].kernel_names.push_back(p.identifier());
            }
        }
        entries[1].arch = arch;
        entries[1].code = hip::compile_to_native(ps.str(), arch);
        hip::write_archive(out_path, entries);
    });
}

int bbfft_aot_generate(const char *const *descriptors, size_t num_descriptors,
                       const char *device_info, const char *arch, const char *out_path,
                       unsigned shard, unsigned num_shards, size_t *num_kernels) {
    return guarded([&] {
        require(descriptors && device_info && arch && out_path, "null argument");
        require(num_shards > 0 && shard < num_shards, "invalid shard");
        auto info = parse_device_info(device_info);
        // Unique kernels in first-seen order; this shard takes every num_shards-th one.
        std::vector<std::pair<std::string, std::string>> kernels; // (name, source)
        std::set<std::string> seen;
        std::size_t index = 0;
        for (size_t i = 0; i < num_descriptors; ++i) {
            auto cfg = parse_fft_descriptor(descriptors[i]);
            auto layout = decompose(cfg);
            // the kernels a plan for this descriptor uses: the fused pair (if any) and the
            // remaining passes
            fused_pair fp;
            bool const fused = find_fused_pair(layout, info, 0u, fp);
            if (fused && seen.insert(fp.fc.identifier()).second) {
                if (index++ % num_shards == shard) {
                    std::ostringstream os;
                    generate_fused2(os, fp.fc);
                    kernels.emplace_back(fp.fc.identifier(), os.str());
                }
            }
            for (std::size_t q = 0; q < layout.passes.size(); ++q) {
                if (fused && (q == fp.first || q == fp.first + 1)) {
                    continue;
                }
                auto kc = configure_fft1d(layout.passes[q], info);
                auto id = kc.identifier();
                if (seen.insert(id).second) {
                    if (index++ % num_shards == shard) {
                        std::ostringstream os;
                        generate_fft1d(os, kc);
                        kernels.emplace_back(id, os.str());
                    }
                }
            }
        }
        // Several kernels per code object keeps the archive and the hiprtc call count small.
        std::size_t const per_object = 16;
        std::vector<hip::archive_entry> entries;
        for (std::size_t k0 = 0; k0 < kernels.size(); k0 += per_object) {
            hip::archive_entry e;
            e.arch = arch;
            std::string src;
            for (std::size_t k = k0; k < std::min(kernels.size(), k0 + per_object); ++k) {
                src += kernels[k].second;
                e.kernel_names.push_back(kernels[k].first);
            }
            e.code = hip::compile_to_native(src, arch);
            entries.push_back(std::move(e));
        }
        hip::write_archive(out_path, entries);
        if (num_kernels) {
            *num_kernels = kernels.size();
        }
    });
}

} // extern "C"
