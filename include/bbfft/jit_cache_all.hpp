// In-memory cache that keeps every module it is given.
// Mirrors the reference interface include/bbfft/jit_cache_all.hpp:20-37.
#ifndef BBFFT_JIT_CACHE_ALL_HPP
#define BBFFT_JIT_CACHE_ALL_HPP

#include "bbfft/export.hpp"
#include "bbfft/jit_cache.hpp"
#include "bbfft/shared_handle.hpp"

#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

namespace bbfft {

/// Thread-safe: get/store take an internal lock (plans may be created from several threads).
class BBFFT_EXPORT jit_cache_all : public jit_cache {
  public:
    auto get(jit_cache_key const &key) const -> shared_handle<module_handle_t> override;
    void store(jit_cache_key const &key, shared_handle<module_handle_t> mod) override;
    auto kernel_names() const -> std::vector<std::string>;

  private:
    mutable std::mutex mtx_;
    std::unordered_map<jit_cache_key, shared_handle<module_handle_t>, jit_cache_key_hash> mods_;
};

} // namespace bbfft

#endif // BBFFT_JIT_CACHE_ALL_HPP
