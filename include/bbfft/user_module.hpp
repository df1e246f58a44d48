// User load/store callbacks inlined into generated kernels.
// Mirrors the reference interface include/bbfft/user_module.hpp:16-37.
#ifndef BBFFT_USER_MODULE_HPP
#define BBFFT_USER_MODULE_HPP

#include "bbfft/export.hpp"

#include <cstddef>

namespace bbfft {

/**
 * @brief Source language of a user module.
 *
 * opencl_c keeps source written for the reference library compiling: the generator prepends a
 * prelude that maps the OpenCL address-space qualifiers and scalar typedefs onto HIP.
 * hip is plain HIP device code.
 */
enum class kernel_language {
    opencl_c, ///< OpenCL-C style callback (compiled through a small HIP prelude)
    hip       ///< HIP device code
};

/**
 * @brief User callbacks.
 *
 * The load function has the signature `T2 load(T2* in, size_t offset)` (or `T load(T* in, size_t
 * offset)` for real input) and the store function `void store(T2* out, size_t offset, T2 value)`
 * (or with T for real output). Offsets are in elements from the tensor base pointer.
 * On 2-D/3-D plans (an extension: the reference accepts callbacks on 1-D plans only) the load
 * callback runs in the plan's first pass and the store callback in its last, with offsets into
 * the plan's input and output tensors; adjoint plans take them where they run in a c2c pass.
 */
struct BBFFT_EXPORT user_module {
    char const *data = nullptr;           ///< Source code string
    std::size_t length = 0;               ///< Length of source code string
    char const *load_function = nullptr;  ///< Name of load function (C string)
    char const *store_function = nullptr; ///< Name of store function (C string)
    kernel_language language = kernel_language::opencl_c; ///< Language

    /// True when source and at least one callback name are set.
    explicit operator bool() const noexcept;
};

} // namespace bbfft

#endif // BBFFT_USER_MODULE_HPP
