// Cache of ahead-of-time compiled kernels.
// Mirrors the reference interface include/bbfft/aot_cache.hpp:21-48.
#ifndef BBFFT_AOT_CACHE_HPP
#define BBFFT_AOT_CACHE_HPP

#include "bbfft/export.hpp"
#include "bbfft/jit_cache.hpp"
#include "bbfft/shared_handle.hpp"

#include <string>
#include <unordered_set>
#include <vector>

namespace bbfft {

/// A loaded native module, the kernel names it contains, and its device id.
struct BBFFT_EXPORT aot_module {
    shared_handle<module_handle_t> mod;
    std::unordered_set<std::string> kernel_names;
    std::uint64_t device_id;
};

/// Read-only cache: get() looks kernels up in the registered modules; store() is a no-op.
class BBFFT_EXPORT aot_cache : public jit_cache {
  public:
    auto get(jit_cache_key const &key) const -> shared_handle<module_handle_t> override;
    void store(jit_cache_key const &key, shared_handle<module_handle_t> mod) override;
    void register_module(aot_module aot_mod);

  private:
    std::vector<aot_module> aot_modules_;
};

} // namespace bbfft

#endif // BBFFT_AOT_CACHE_HPP
