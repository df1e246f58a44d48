// HIP plans: module lookup/JIT, launch geometry and event handling.
//
// Reference behaviour: src/common/algorithm.hpp:20-30 (dimension check and dispatch),
// src/common/algorithm/small_batch_fft.hpp:42-139 and factor2_slm_fft.hpp:92-203 (cache lookup,
// generate on miss, in-place guard, bind (in, out, K) and launch), nd_fft.hpp:114-176 (chained
// passes through a scratch buffer). One kernel family replaces the two reference families, the
// kernel arguments are (in, out, K) and the launch grid is linearised over (M tile, K tile).
#include "bbfft/hip/make_plan.hpp"
#include "../plan/decompose.hpp"
#include "bbfft/bad_configuration.hpp"
#include "bbfft/detail/generator_impl.hpp"
#include "bbfft/hip/device.hpp"
#include "bbfft/hip/error.hpp"
#include "bbfft/hip/online_compiler.hpp"
#include "hip_plan_impl.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <sstream>

namespace bbfft {

namespace {

void release_module(module_handle_t m) { (void)hipModuleUnload(reinterpret_cast<hipModule_t>(m)); }

} // namespace

namespace hip_detail {

fft1d_plan::fft1d_plan(configuration const &cfg, hipStream_t stream, int device,
                       jit_cache *cache, fft1d_tuning const &tuning, unsigned variant)
    : stream_(stream), device_(device) {
    kc_ = configure_fft1d(cfg, hip::get_device_info(device), tuning, variant);
    name_ = kc_.identifier();
    K_ = cfg.shape[2];
    jit_cache_key key{name_, hip::get_device_id(device)};
    if (cache) {
        mod_ = cache->get(key);
    }
    if (!mod_) {
        std::ostringstream src;
        generate_fft1d(src, kc_, {}, &cfg.callbacks);
        hipModule_t m = hip::build_kernel_bundle(src.str(), device);
        mod_ = shared_handle<module_handle_t>(reinterpret_cast<module_handle_t>(m),
                                              &release_module);
        if (cache) {
            cache->store(key, mod_);
        }
    }
    fn_ = hip::create_kernel(reinterpret_cast<hipModule_t>(mod_.get()), name_);
    std::size_t const mg = (kc_.M + kc_.Mb - 1) / kc_.Mb;
    std::size_t const kcols = kc_.K_columns(K_);
    std::size_t const ntiles = mg * ((kcols + kc_.Kb - 1) / kc_.Kb);
    grid_ = static_cast<unsigned>(std::min<std::size_t>(ntiles, std::size_t(1) << 24));
    if (kc_.groups_per_cu) {
        int cus = 0;
        BBFFT_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        grid_ = static_cast<unsigned>(std::max<std::size_t>(
            1, std::min<std::size_t>(ntiles, std::size_t(kc_.groups_per_cu) *
                                                 static_cast<std::size_t>(std::max(cus, 1)))));
    }
    block_ = static_cast<unsigned>(kc_.work_group_size());
    std::size_t const rb = static_cast<std::size_t>(kc_.fp);
    // Real data read or written as (x[2n], x[2n+1]) pairs needs complex alignment.
    in_align_ = kc_.type == transform_type::r2c ? rb : 2 * rb;
    if (kc_.flags & (fft1d_input_f16 | fft1d_input_bf16)) {
        in_align_ = 4; // (re, im) half pairs
    }
    out_align_ = kc_.type == transform_type::c2r ? rb : 2 * rb;
    if (kc_.N % 2 == 0 && kc_.M == 1 && kc_.type == transform_type::r2c &&
        !kc_.load_function && kc_.istride[1] == 1 && kc_.istride[2] % 2 == 0) {
        in_align_ = 2 * rb;
    }
    if (kc_.N % 2 == 0 && kc_.M == 1 && kc_.type == transform_type::c2r &&
        !kc_.store_function && kc_.ostride[1] == 1 && kc_.ostride[2] % 2 == 0) {
        out_align_ = 2 * rb;
    }
}

void fft1d_plan::launch(void const *in, void *out) {
    if (in == out && kc_.inplace_unsupported) {
        throw bad_configuration("The plan does not support in-place transform on the current "
                                "device. Please use the out-of-place transform.");
    }
    if (reinterpret_cast<std::uintptr_t>(in) % in_align_ ||
        reinterpret_cast<std::uintptr_t>(out) % out_align_) {
        throw bad_configuration("Input/output pointers are not aligned to the element size");
    }
    void const *a0 = in;
    void *a1 = out;
    unsigned long k = K_;
    void *args[] = {&a0, &a1, &k};
    BBFFT_HIP_CHECK(
        hipModuleLaunchKernel(fn_, grid_, 1, 1, block_, 1, 1, 0, stream_, args, nullptr));
}

void fft1d_plan::execute(void const *in, void *out, hipEvent_t signal, std::uint32_t nwait,
                         hipEvent_t *wait) {
    for (std::uint32_t i = 0; i < nwait; ++i) {
        BBFFT_HIP_CHECK(hipStreamWaitEvent(stream_, wait[i], 0));
    }
    launch(in, out);
    if (signal) {
        BBFFT_HIP_CHECK(hipEventRecord(signal, stream_));
    }
}

std::vector<launch_info> fft1d_plan::launches() const {
    return {launch_info{name_, grid_, block_, kc_.lds_bytes()}};
}

fused2_kernel::fused2_kernel(fused2_configuration const &fc, hipStream_t stream, int device,
                             jit_cache *cache)
    : stream_(stream), fc_(fc) {
    name_ = fc_.identifier();
    jit_cache_key key{name_, hip::get_device_id(device)};
    if (cache) {
        mod_ = cache->get(key);
    }
    if (!mod_) {
        std::ostringstream src;
        generate_fused2(src, fc_);
        hipModule_t m = hip::build_kernel_bundle(src.str(), device);
        mod_ = shared_handle<module_handle_t>(reinterpret_cast<module_handle_t>(m),
                                              &release_module);
        if (cache) {
            cache->store(key, mod_);
        }
    }
    fn_ = hip::create_kernel(reinterpret_cast<hipModule_t>(mod_.get()), name_);
    grid_ = static_cast<unsigned>(std::min<std::size_t>(fc_.planes, std::size_t(1) << 24));
    if (fc_.stage_words) {
        // staged kernels overlap planes inside a work-group: one resident group per slot
        int per_cu = 0, cus = 0;
        BBFFT_HIP_CHECK(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, fn_, static_cast<int>(fc_.group_size), 0));
        BBFFT_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                              device));
        std::size_t const resident =
            static_cast<std::size_t>(std::max(per_cu, 1)) * static_cast<std::size_t>(std::max(cus, 1));
        grid_ = static_cast<unsigned>(std::min<std::size_t>(fc_.planes, resident));
    }
}

void fused2_kernel::launch(void const *in, void *out) {
    std::size_t const rb = static_cast<std::size_t>(fc_.a.fp);
    std::size_t const ia = fc_.a.type == transform_type::r2c ? rb : 2 * rb;
    std::size_t const oa = fc_.b.type == transform_type::c2r ? rb : 2 * rb;
    if (reinterpret_cast<std::uintptr_t>(in) % ia || reinterpret_cast<std::uintptr_t>(out) % oa) {
        throw bad_configuration("Input/output pointers are not aligned to the element size");
    }
    void const *a0 = in;
    void *a1 = out;
    unsigned long planes = fc_.planes;
    void *args[] = {&a0, &a1, &planes};
    BBFFT_HIP_CHECK(hipModuleLaunchKernel(fn_, grid_, 1, 1,
                                          static_cast<unsigned>(fc_.group_size), 1, 1, 0,
                                          stream_, args, nullptr));
}

launch_info fused2_kernel::info() const {
    return launch_info{name_, grid_, static_cast<unsigned>(fc_.group_size),
                       fc_.lds_bytes() + 4 * fc_.stage_words};
}

nd_plan::nd_plan(pass_layout const &layout, hipStream_t stream, int device, jit_cache *cache,
                 unsigned real_pass_variant)
    : stream_(stream), device_(device) {
    fused_pair fp;
    if (find_fused_pair(layout, hip::get_device_info(device), real_pass_variant, fp)) {
        fused_first_ = fp.first;
        fused_ = std::make_unique<fused2_kernel>(fp.fc, stream, device, cache);
    }
    for (std::size_t i = 0; i < layout.passes.size(); ++i) {
        auto const &p = layout.passes[i];
        if (fused_ && (i == fused_first_ || i == fused_first_ + 1)) {
            passes_.push_back(nullptr);
            continue;
        }
        unsigned v = p.type == transform_type::c2c ? 0u : real_pass_variant;
        passes_.push_back(std::make_unique<fft1d_plan>(p, stream, device, cache,
                                                       fft1d_tuning{}, v));
    }
    if (layout.tmp_bytes) {
        BBFFT_HIP_CHECK(hipMalloc(&tmp_, layout.tmp_bytes));
        tmp_bytes_ = layout.tmp_bytes;
    }
}

nd_plan::~nd_plan() {
    if (tmp_) {
        (void)hipFree(tmp_);
    }
}

void nd_plan::execute(void const *in, void *out, hipEvent_t signal, std::uint32_t nwait,
                      hipEvent_t *wait) {
    for (std::uint32_t i = 0; i < nwait; ++i) {
        BBFFT_HIP_CHECK(hipStreamWaitEvent(stream_, wait[i], 0));
    }
    run(in, out, tmp_);
    if (signal) {
        BBFFT_HIP_CHECK(hipEventRecord(signal, stream_));
    }
}

void nd_plan::execute_with_scratch(void const *in, void *out, void *scratch) {
    if (tmp_bytes_ && !scratch) {
        throw bad_configuration("The plan needs a scratch buffer");
    }
    run(in, out, tmp_bytes_ ? scratch : nullptr);
}

void nd_plan::run(void const *in, void *out, void *tmp) {
    // All steps run on one stream, so stream order replaces the reference's internal events.
    // Step 0 reads `in`, the last step writes `out`, the others work in the scratch buffer (or
    // in `out` when the input is not larger than the output).
    if (!tmp) {
        tmp = out;
    }
    std::size_t const D = passes_.size();
    std::size_t const steps = fused_ ? D - 1 : D;
    std::size_t step = 0;
    for (std::size_t i = 0; i < D; ++i, ++step) {
        void const *src = step == 0 ? in : tmp;
        void *dst = step + 1 == steps ? out : tmp;
        if (fused_ && i == fused_first_) {
            fused_->launch(src, dst);
            ++i;
        } else {
            passes_[i]->launch(src, dst);
        }
    }
}

std::vector<launch_info> nd_plan::launches() const {
    std::vector<launch_info> all;
    for (std::size_t i = 0; i < passes_.size(); ++i) {
        if (fused_ && i == fused_first_) {
            all.push_back(fused_->info());
            ++i;
            continue;
        }
        auto l = passes_[i]->launches();
        all.insert(all.end(), l.begin(), l.end());
    }
    return all;
}

std::shared_ptr<plan_base> create(configuration const &cfg, hipStream_t stream,
                                  jit_cache *cache) {
    if (cfg.dim < 1 || cfg.dim > max_fft_dim) {
        throw bad_configuration("Unsupported FFT dimension: " + std::to_string(cfg.dim));
    }
    // Host-side validation first: configuration errors never depend on the device.
    auto layout = decompose(cfg);
    int device = 0;
    BBFFT_HIP_CHECK(hipGetDevice(&device));
    if (layout.passes.size() == 1) {
        return std::make_shared<fft1d_plan>(layout.passes[0], stream, device, cache);
    }
    return std::make_shared<nd_plan>(layout, stream, device, cache);
}

std::shared_ptr<plan_base> create_adjoint(configuration const &cfg, hipStream_t stream,
                                          jit_cache *cache) {
    if (cfg.dim < 1 || cfg.dim > max_fft_dim) {
        throw bad_configuration("Unsupported FFT dimension: " + std::to_string(cfg.dim));
    }
    if (cfg.callbacks && (cfg.dim == 1 && cfg.type != transform_type::c2c)) {
        // the real pass of a 1-D real adjoint carries the fold weights; callbacks are taken
        // where they run in a c2c pass (c2c plans, and the first or last pass of 2-D/3-D plans)
        throw bad_configuration("Adjoint plans of 1-D real transforms do not support callbacks");
    }
    unsigned variant = 0;
    auto adj = adjoint_configuration(cfg, &variant);
    auto layout = decompose(adj);
    for (auto const &p : layout.passes) {
        if (p.callbacks && p.type != transform_type::c2c) {
            throw bad_configuration("Adjoint plans take callbacks in c2c passes only");
        }
    }
    int device = 0;
    BBFFT_HIP_CHECK(hipGetDevice(&device));
    if (layout.passes.size() == 1) {
        return std::make_shared<fft1d_plan>(layout.passes[0], stream, device, cache,
                                            fft1d_tuning{}, variant);
    }
    return std::make_shared<nd_plan>(layout, stream, device, cache, variant);
}

std::shared_ptr<plan_base> create_variant(configuration const &cfg, unsigned variant,
                                          hipStream_t stream, jit_cache *cache) {
    auto layout = decompose(cfg);
    if (layout.passes.size() != 1) {
        throw bad_configuration("Kernel variants apply to 1-D configurations only");
    }
    int device = 0;
    BBFFT_HIP_CHECK(hipGetDevice(&device));
    return std::make_shared<fft1d_plan>(layout.passes[0], stream, device, cache, fft1d_tuning{},
                                        variant);
}

std::shared_ptr<plan_base> create_tuned(configuration const &cfg, fft1d_tuning const &tuning,
                                        hipStream_t stream, jit_cache *cache) {
    auto layout = decompose(cfg);
    if (layout.passes.size() != 1) {
        throw bad_configuration("Explicit tuning applies to 1-D configurations only");
    }
    int device = 0;
    BBFFT_HIP_CHECK(hipGetDevice(&device));
    return std::make_shared<fft1d_plan>(layout.passes[0], stream, device, cache, tuning);
}

} // namespace hip_detail

auto make_plan(configuration const &cfg, hipStream_t stream, jit_cache *cache) -> hip_plan {
    return hip_plan(hip_detail::create(cfg, stream, cache));
}

auto make_half_input_plan(configuration const &cfg, unsigned flags, hipStream_t stream,
                          jit_cache *cache) -> hip_plan {
    unsigned const allowed = fft1d_input_f16 | fft1d_input_bf16 | fft1d_mfma;
    if ((flags & ~allowed) || !(flags & (fft1d_input_f16 | fft1d_input_bf16))) {
        throw bad_configuration("Half-input plans take fft1d_input_f16 or fft1d_input_bf16, "
                                "optionally with fft1d_mfma");
    }
    return hip_plan(hip_detail::create_variant(cfg, flags, stream, cache));
}

} // namespace bbfft
