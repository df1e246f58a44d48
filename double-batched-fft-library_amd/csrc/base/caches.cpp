// jit_cache key hashing, the keep-everything cache and the AOT lookup cache.
// Reference behaviour: src/base/jit_cache.cpp:10-19, src/base/jit_cache_all.cpp:10-26,
// src/base/aot_cache.cpp:12-28.
#include "bbfft/aot_cache.hpp"
#include "bbfft/jit_cache.hpp"
#include "bbfft/jit_cache_all.hpp"

#include <functional>
#include <utility>

namespace bbfft {

bool jit_cache_key::operator==(jit_cache_key const &other) const {
    return device_id == other.device_id && kernel_name == other.kernel_name;
}

std::size_t jit_cache_key_hash::operator()(jit_cache_key const &key) const noexcept {
    // 64-bit mix of the two member hashes.
    std::uint64_t h = std::hash<std::string>{}(key.kernel_name);
    std::uint64_t d = key.device_id * 0x9e3779b97f4a7c15ull;
    h ^= d + 0x7f4a7c159e3779b9ull + (h << 12) + (h >> 4);
    return static_cast<std::size_t>(h);
}

jit_cache::~jit_cache() {}

auto jit_cache_all::get(jit_cache_key const &key) const -> shared_handle<module_handle_t> {
    std::lock_guard<std::mutex> lock(mtx_);
    auto it = mods_.find(key);
    return it == mods_.end() ? shared_handle<module_handle_t>{} : it->second;
}

void jit_cache_all::store(jit_cache_key const &key, shared_handle<module_handle_t> mod) {
    std::lock_guard<std::mutex> lock(mtx_);
    mods_[key] = std::move(mod);
}

auto jit_cache_all::kernel_names() const -> std::vector<std::string> {
    std::lock_guard<std::mutex> lock(mtx_);
    std::vector<std::string> names;
    names.reserve(mods_.size());
    for (auto const &kv : mods_) {
        names.push_back(kv.first.kernel_name);
    }
    return names;
}

auto aot_cache::get(jit_cache_key const &key) const -> shared_handle<module_handle_t> {
    for (auto const &m : aot_modules_) {
        if (m.device_id == key.device_id && m.kernel_names.count(key.kernel_name)) {
            return m.mod;
        }
    }
    return {};
}

void aot_cache::store(jit_cache_key const &, shared_handle<module_handle_t>) {}

void aot_cache::register_module(aot_module aot_mod) { aot_modules_.push_back(std::move(aot_mod)); }

} // namespace bbfft
