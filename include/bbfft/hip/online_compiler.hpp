// hiprtc compilation and code-object loading.
// Plays the role of the reference's include/bbfft/ze/online_compiler.hpp (ocloc) and
// include/bbfft/sycl/online_compiler.hpp.
#ifndef BBFFT_HIP_ONLINE_COMPILER_HPP
#define BBFFT_HIP_ONLINE_COMPILER_HPP

#include "bbfft/aot_cache.hpp"
#include "bbfft/export.hpp"
#include "bbfft/module_format.hpp"

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <string>
#include <vector>

namespace bbfft::hip {

/// Compile HIP source to a native code object for `arch` (e.g. "gfx950"). No device needed.
BBFFT_EXPORT std::vector<std::uint8_t> compile_to_native(std::string const &source,
                                                         std::string const &arch,
                                                         std::vector<std::string> const &options = {});
/// Compile and load a module on the current device.
BBFFT_EXPORT hipModule_t build_kernel_bundle(std::string const &source, int device,
                                             std::vector<std::string> const &options = {});
/// Load a module from a native code object.
BBFFT_EXPORT hipModule_t build_kernel_bundle(std::uint8_t const *binary, std::size_t binary_size,
                                             module_format format, int device);
/// Get a kernel from a module.
BBFFT_EXPORT hipFunction_t create_kernel(hipModule_t mod, std::string const &name);
/// Load a code object holding the named kernels as an aot_module.
BBFFT_EXPORT aot_module create_aot_module(std::uint8_t const *binary, std::size_t binary_size,
                                          module_format format,
                                          std::vector<std::string> const &kernel_names,
                                          int device);
/**
 * @brief Register every code object of an AOT archive file with `cache`.
 *
 * The archive (written by bbfft-aot-generate) holds code objects and their kernel names.
 * Returns the number of kernels registered; throws bbfft::hip::error if the file is unreadable
 * or its architecture differs from the device's.
 */
BBFFT_EXPORT std::size_t register_aot_archive(aot_cache &cache, std::string const &path,
                                              int device);
/// Same for an archive held in memory (e.g. linked into the executable, see
/// `python -m bbfft_amd.aot --embed`; the reference links archives with `ld -r -b binary`,
/// cmake/AddAotKernelsToTarget.cmake:4-33).
BBFFT_EXPORT std::size_t register_aot_archive(aot_cache &cache, std::uint8_t const *data,
                                              std::size_t size, int device);

} // namespace bbfft::hip

#endif // BBFFT_HIP_ONLINE_COMPILER_HPP
