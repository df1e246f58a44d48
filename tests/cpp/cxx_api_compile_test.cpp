// Compile-time check of the C++ drop-in API: the calls a reference user makes, against the
// gfx950 headers (make_plan with a HIP stream instead of a SYCL queue). Linked, not run.
#include "bbfft/aot_cache.hpp"
#include "bbfft/bad_configuration.hpp"
#include "bbfft/configuration.hpp"
#include "bbfft/detail/generator_impl.hpp"
#include "bbfft/generator.hpp"
#include "bbfft/hip/device.hpp"
#include "bbfft/hip/make_plan.hpp"
#include "bbfft/hip/online_compiler.hpp"
#include "bbfft/jit_cache_all.hpp"
#include "bbfft/parser.hpp"
#include "bbfft/tensor_indexer.hpp"

#include <sstream>

int run(hipStream_t stream, void *x, void *y) {
    using namespace bbfft;
    configuration cfg = {1, {8, 256, 32768}, precision::f32, direction::forward};
    cfg.set_strides_default(false);
    jit_cache_all cache;
    auto plan = make_plan(cfg, stream, &cache);
    plan.execute(x, y);
    auto cfg2 = parse_fft_descriptor("srfi128x128x128*32");
    auto plan2 = make_plan(cfg2, stream);
    plan2.execute(x);
    std::ostringstream os;
    auto names = generate_fft_kernels(os, {cfg, cfg2}, parse_device_info("{1024, {64}, 163840, gpu}"));
    aot_cache aot;
    hip::register_aot_archive(aot, "kernels.bin", 0);
    // extensions: ragged batches and half-precision input with the MFMA stage
    ragged_configuration rc{precision::f32, {ragged_job{2, 60, 100, direction::forward, 0, 0}}};
    auto plan3 = make_ragged_plan(rc, stream, &cache, true);
    plan3.execute(x, y);
    auto plan4 = make_half_input_plan(parse_fft_descriptor("scfo4.512*64"),
                                      fft1d_input_f16 | fft1d_mfma, stream, &cache);
    plan4.execute(x, y);
    return static_cast<int>(names.size()) + static_cast<int>(plan ? 1 : 0);
}
