// Ragged plans: one persistent kernel launch for a list of c2c jobs of different shapes.
//
// No counterpart in the reference. The job table is uploaded once at plan creation; a launch
// takes (in, out, table, njobs) and a grid of 16 work-groups per resident slot of the device.
// With use_graph the launch is captured into a HIP graph once per (in, out) pair and replayed
// with hipGraphLaunch.
#include "bbfft/bad_configuration.hpp"
#include "bbfft/detail/generator_impl.hpp"
#include "bbfft/hip/device.hpp"
#include "bbfft/hip/error.hpp"
#include "bbfft/hip/make_plan.hpp"
#include "bbfft/hip/online_compiler.hpp"
#include "hip_plan_impl.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdint>
#include <sstream>

namespace bbfft {

namespace hip_detail {

namespace {
void release_ragged_module(module_handle_t m) {
    (void)hipModuleUnload(reinterpret_cast<hipModule_t>(m));
}
} // namespace

ragged_plan::ragged_plan(ragged_configuration const &cfg, hipStream_t stream, int device,
                         jit_cache *cache, bool use_graph)
    : stream_(stream), device_(device), use_graph_(use_graph) {
    rk_ = configure_ragged(cfg, hip::get_device_info(device));
    name_ = rk_.identifier();
    jit_cache_key key{name_, hip::get_device_id(device)};
    if (cache) {
        mod_ = cache->get(key);
    }
    if (!mod_) {
        std::ostringstream src;
        generate_ragged(src, rk_);
        hipModule_t m = hip::build_kernel_bundle(src.str(), device);
        mod_ = shared_handle<module_handle_t>(reinterpret_cast<module_handle_t>(m),
                                              &release_ragged_module);
        if (cache) {
            cache->store(key, mod_);
        }
    }
    fn_ = hip::create_kernel(reinterpret_cast<hipModule_t>(mod_.get()), name_);
    block_ = static_cast<unsigned>(rk_.group_size);
    int per_cu = 0;
    BBFFT_HIP_CHECK(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn_,
                                                                      static_cast<int>(block_),
                                                                      0));
    int cus = 0;
    BBFFT_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    std::uint64_t const resident =
        static_cast<std::uint64_t>(std::max(per_cu, 1)) * static_cast<std::uint64_t>(std::max(cus, 1));
    // 16 work-groups per resident slot, each with its own contiguous range of tiles: the
    // dispatcher refills slots as ranges finish, which evens out the end of the launch. On
    // MI355X (4096-job serving mix) 16 per slot took 67.8 ms against 78.9 ms with one per slot,
    // 68.6-69.5 ms with 8 or 64-128 per slot and 73.8 ms with one work-group per tile.
    constexpr std::uint64_t groups_per_slot = 16;
    grid_ = static_cast<unsigned>(std::max<std::uint64_t>(
        1, std::min<std::uint64_t>(rk_.total_tiles(), resident * groups_per_slot)));
    for (auto const &job : cfg.jobs) {
        std::size_t const n = job.M * job.N * job.K;
        in_end_ = std::max(in_end_, job.in_offset + n);
        out_end_ = std::max(out_end_, job.out_offset + n);
    }
    auto table = ragged_job_table(cfg, rk_, grid_);
    njobs_ = cfg.jobs.size();
    BBFFT_HIP_CHECK(hipMalloc(&table_, table.size() * sizeof(std::uint64_t)));
    hipError_t const copied = hipMemcpy(table_, table.data(),
                                        table.size() * sizeof(std::uint64_t),
                                        hipMemcpyHostToDevice);
    if (copied != hipSuccess) {
        // a throwing constructor runs no destructor: release the table here
        (void)hipFree(table_);
        table_ = nullptr;
        BBFFT_HIP_CHECK(copied);
    }
}

ragged_plan::~ragged_plan() {
    release_graph();
    if (table_) {
        (void)hipFree(table_);
    }
}

void ragged_plan::release_graph() {
    if (exec_) {
        (void)hipGraphExecDestroy(exec_);
        exec_ = nullptr;
    }
}

void ragged_plan::launch(void const *in, void *out, hipStream_t stream) {
    void const *a0 = in;
    void *a1 = out;
    void const *a2 = table_;
    unsigned long n = njobs_;
    void *args[] = {&a0, &a1, &a2, &n};
    BBFFT_HIP_CHECK(
        hipModuleLaunchKernel(fn_, grid_, 1, 1, block_, 1, 1, 0, stream, args, nullptr));
}

void ragged_plan::execute(void const *in, void *out, hipEvent_t signal, std::uint32_t nwait,
                          hipEvent_t *wait) {
    std::size_t const align = 2 * static_cast<std::size_t>(rk_.fp);
    if (reinterpret_cast<std::uintptr_t>(in) % align ||
        reinterpret_cast<std::uintptr_t>(out) % align) {
        throw bad_configuration("Input/output pointers are not aligned to the element size");
    }
    for (std::uint32_t i = 0; i < nwait; ++i) {
        BBFFT_HIP_CHECK(hipStreamWaitEvent(stream_, wait[i], 0));
    }
    if (use_graph_) {
        if (!exec_ || in != graph_in_ || out != graph_out_) {
            release_graph();
            // Capture on a private stream (the plan's stream may be the null stream, which
            // cannot be captured); the instantiated graph is launched on the plan's stream.
            hipStream_t cs = nullptr;
            BBFFT_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
            hipGraph_t graph = nullptr;
            hipError_t err = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
            if (err == hipSuccess) {
                try {
                    launch(in, out, cs);
                } catch (...) {
                    (void)hipStreamEndCapture(cs, &graph);
                    if (graph) {
                        (void)hipGraphDestroy(graph);
                    }
                    (void)hipStreamDestroy(cs);
                    throw;
                }
                err = hipStreamEndCapture(cs, &graph);
            }
            if (err == hipSuccess) {
                err = hipGraphInstantiate(&exec_, graph, nullptr, nullptr, 0);
            }
            if (graph) {
                (void)hipGraphDestroy(graph);
            }
            (void)hipStreamDestroy(cs);
            BBFFT_HIP_CHECK(err);
            graph_in_ = in;
            graph_out_ = out;
        }
        BBFFT_HIP_CHECK(hipGraphLaunch(exec_, stream_));
    } else {
        launch(in, out, stream_);
    }
    if (signal) {
        BBFFT_HIP_CHECK(hipEventRecord(signal, stream_));
    }
}

void ragged_plan::execute_with_scratch(void const *in, void *out, void *) {
    std::size_t const align = 2 * static_cast<std::size_t>(rk_.fp);
    if (reinterpret_cast<std::uintptr_t>(in) % align ||
        reinterpret_cast<std::uintptr_t>(out) % align) {
        throw bad_configuration("Input/output pointers are not aligned to the element size");
    }
    launch(in, out, stream_);
}

std::vector<launch_info> ragged_plan::launches() const {
    return {launch_info{name_, grid_, block_, rk_.lds_bytes()}};
}

} // namespace hip_detail

auto make_ragged_plan(ragged_configuration const &cfg, hipStream_t stream, jit_cache *cache,
                      bool use_graph) -> hip_plan {
    int device = 0;
    BBFFT_HIP_CHECK(hipGetDevice(&device));
    return hip_plan(
        std::make_shared<hip_detail::ragged_plan>(cfg, stream, device, cache, use_graph));
}

} // namespace bbfft
