// Reference-counted opaque handle (module handles shared between plans and caches).
// Mirrors the reference interface include/bbfft/shared_handle.hpp:18-52.
#ifndef BBFFT_SHARED_HANDLE_HPP
#define BBFFT_SHARED_HANDLE_HPP

#include <cstdint>
#include <memory>
#include <type_traits>

namespace bbfft {

template <typename T, typename Enable = void> class shared_handle;

/// Shared ownership of a pointer-sized handle with a custom release function.
template <typename T>
class shared_handle<T,
                    std::enable_if_t<sizeof(T) <= sizeof(void *) && alignof(T) <= sizeof(void *)>> {
  public:
    shared_handle() = default;
    shared_handle(T t, void (*release)(T))
        : ref_(reinterpret_cast<void *>(t), [release](void *p) { release(reinterpret_cast<T>(p)); }) {}
    T get() const { return reinterpret_cast<T>(ref_.get()); }
    explicit operator bool() const noexcept { return static_cast<bool>(ref_); }

  private:
    std::shared_ptr<void> ref_;
};

} // namespace bbfft

#endif // BBFFT_SHARED_HANDLE_HPP
