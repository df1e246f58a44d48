// Internal plan classes of the HIP back end (shared by make_plan and the C ABI).
#ifndef BBFFT_AMD_HIP_PLAN_IMPL_HPP
#define BBFFT_AMD_HIP_PLAN_IMPL_HPP

#include "../plan/decompose.hpp"
#include "bbfft/configuration.hpp"
#include "bbfft/detail/generator_impl.hpp"
#include "bbfft/detail/plan_impl.hpp"
#include "bbfft/jit_cache.hpp"
#include "bbfft/ragged.hpp"
#include "bbfft/shared_handle.hpp"

#include <hip/hip_runtime_api.h>

#include <memory>
#include <string>
#include <vector>

namespace bbfft::hip_detail {

/// One kernel launch of a plan (for introspection, profiling and tests).
struct launch_info {
    std::string kernel_name;
    unsigned grid;
    unsigned block;
    std::size_t lds_bytes;
};

class plan_base : public detail::plan_unmanaged_event_impl<hipEvent_t> {
  public:
    virtual std::vector<launch_info> launches() const = 0;
    /// Same as execute (without events) with `scratch` (at least scratch_bytes() bytes of
    /// device memory) in place of the plan's own scratch buffer; allocates nothing.
    virtual void execute_with_scratch(void const *in, void *out, void *scratch) = 0;
    /// Device memory the plan needs between its passes (0: none).
    virtual std::size_t scratch_bytes() const { return 0; }
    virtual hipStream_t stream() const = 0;
    virtual int device() const = 0;
};

/// A single 1-D double-batched kernel.
class fft1d_plan final : public plan_base {
  public:
    fft1d_plan(configuration const &cfg, hipStream_t stream, int device, jit_cache *cache,
               fft1d_tuning const &tuning = {}, unsigned variant = 0u);
    void execute(void const *in, void *out, hipEvent_t signal, std::uint32_t nwait,
                 hipEvent_t *wait) override;
    /// Launch on the plan's stream without event handling.
    void launch(void const *in, void *out);
    void execute_with_scratch(void const *in, void *out, void *) override { launch(in, out); }
    hipStream_t stream() const override { return stream_; }
    int device() const override { return device_; }
    std::vector<launch_info> launches() const override;
    fft1d_configuration const &kernel_config() const { return kc_; }

  private:
    hipStream_t stream_;
    int device_;
    fft1d_configuration kc_;
    std::string name_;
    shared_handle<module_handle_t> mod_;
    hipFunction_t fn_ = nullptr;
    unsigned long K_ = 0;
    unsigned grid_ = 0, block_ = 0;
    std::size_t in_align_ = 1, out_align_ = 1;
};

/// Two passes of a 2-D/3-D plan in one kernel (planes kept in LDS).
class fused2_kernel {
  public:
    fused2_kernel(fused2_configuration const &fc, hipStream_t stream, int device,
                  jit_cache *cache);
    void launch(void const *in, void *out);
    launch_info info() const;

  private:
    hipStream_t stream_;
    fused2_configuration fc_;
    std::string name_;
    shared_handle<module_handle_t> mod_;
    hipFunction_t fn_ = nullptr;
    unsigned grid_ = 0;
};

/// 2-D/3-D transform as a chain of 1-D passes on one stream; two of them may run fused.
class nd_plan final : public plan_base {
  public:
    nd_plan(pass_layout const &layout, hipStream_t stream, int device, jit_cache *cache,
            unsigned real_pass_variant = 0u);
    ~nd_plan();
    nd_plan(nd_plan const &) = delete; // owns the scratch buffer
    nd_plan &operator=(nd_plan const &) = delete;
    void execute(void const *in, void *out, hipEvent_t signal, std::uint32_t nwait,
                 hipEvent_t *wait) override;
    void execute_with_scratch(void const *in, void *out, void *scratch) override;
    std::size_t scratch_bytes() const override { return tmp_bytes_; }
    hipStream_t stream() const override { return stream_; }
    int device() const override { return device_; }
    std::vector<launch_info> launches() const override;

  private:
    void run(void const *in, void *out, void *tmp);

    hipStream_t stream_;
    int device_;
    std::vector<std::unique_ptr<fft1d_plan>> passes_; ///< null where a fused kernel runs
    std::unique_ptr<fused2_kernel> fused_;
    std::size_t fused_first_ = static_cast<std::size_t>(-1);
    void *tmp_ = nullptr;
    std::size_t tmp_bytes_ = 0;
};

/// Ragged batch: one persistent kernel over a device job table, optionally replayed from a
/// HIP graph that holds the launch.
class ragged_plan final : public plan_base {
  public:
    ragged_plan(ragged_configuration const &cfg, hipStream_t stream, int device,
                jit_cache *cache, bool use_graph);
    ~ragged_plan();
    ragged_plan(ragged_plan const &) = delete;
    ragged_plan &operator=(ragged_plan const &) = delete;
    void execute(void const *in, void *out, hipEvent_t signal, std::uint32_t nwait,
                 hipEvent_t *wait) override;
    /// Direct launch (no graph capture, which may allocate).
    void execute_with_scratch(void const *in, void *out, void *) override;
    hipStream_t stream() const override { return stream_; }
    int device() const override { return device_; }
    std::vector<launch_info> launches() const override;
    ragged_kernel_configuration const &kernel_config() const { return rk_; }

  private:
    void launch(void const *in, void *out, hipStream_t stream);
    void release_graph();

    hipStream_t stream_;
    int device_;
    ragged_kernel_configuration rk_;
    std::string name_;
    shared_handle<module_handle_t> mod_;
    hipFunction_t fn_ = nullptr;
    void *table_ = nullptr;
    unsigned long njobs_ = 0;
    unsigned grid_ = 0, block_ = 0;
    std::size_t in_end_ = 0, out_end_ = 0; ///< complex elements touched from in / out
    bool use_graph_ = false;
    hipGraphExec_t exec_ = nullptr;
    void const *graph_in_ = nullptr;
    void *graph_out_ = nullptr;
};

std::shared_ptr<plan_base> create(configuration const &cfg, hipStream_t stream,
                                  jit_cache *cache);
/// Plan of the adjoint of cfg's transform (see adjoint_configuration).
std::shared_ptr<plan_base> create_adjoint(configuration const &cfg, hipStream_t stream,
                                          jit_cache *cache);
/// 1-D plan of a kernel variant (fft1d_flag bits, e.g. half-precision input and MFMA stage).
std::shared_ptr<plan_base> create_variant(configuration const &cfg, unsigned variant,
                                          hipStream_t stream, jit_cache *cache);
/// 1-D plan with explicit tuning.
std::shared_ptr<plan_base> create_tuned(configuration const &cfg, fft1d_tuning const &tuning,
                                        hipStream_t stream, jit_cache *cache);

} // namespace bbfft::hip_detail

#endif // BBFFT_AMD_HIP_PLAN_IMPL_HPP
