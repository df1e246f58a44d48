// Kernel module cache interface keyed by (kernel identifier, device id).
// Mirrors the reference interface include/bbfft/jit_cache.hpp:17-69.
#ifndef BBFFT_JIT_CACHE_HPP
#define BBFFT_JIT_CACHE_HPP

#include "bbfft/export.hpp"
#include "bbfft/shared_handle.hpp"

#include <cstddef>
#include <cstdint>
#include <limits>
#include <string>

namespace bbfft {

using device_handle_t = std::uintptr_t; ///< opaque device handle
using module_handle_t = std::uintptr_t; ///< opaque module handle (hipModule_t on HIP)

/// Cache key: kernel identifier (= kernel symbol name) and device id.
struct BBFFT_EXPORT jit_cache_key {
    std::string kernel_name = {};
    std::uint64_t device_id = std::numeric_limits<std::uint64_t>::max();
    bool operator==(jit_cache_key const &other) const;
};

struct BBFFT_EXPORT jit_cache_key_hash {
    std::size_t operator()(jit_cache_key const &key) const noexcept;
};

/// Abstract module cache.
class BBFFT_EXPORT jit_cache {
  public:
    virtual ~jit_cache();
    /// Returns an empty handle on a miss.
    virtual auto get(jit_cache_key const &key) const -> shared_handle<module_handle_t> = 0;
    virtual void store(jit_cache_key const &key, shared_handle<module_handle_t> mod) = 0;
};

} // namespace bbfft

#endif // BBFFT_JIT_CACHE_HPP
