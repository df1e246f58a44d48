// Binary module formats accepted by the module loaders.
// Mirrors the reference interface include/bbfft/module_format.hpp:12-15.
#ifndef BBFFT_MODULE_FORMAT_HPP
#define BBFFT_MODULE_FORMAT_HPP

namespace bbfft {

enum class module_format {
    spirv, ///< not produced for gfx950; kept for source compatibility
    native ///< AMDGPU code object (ELF) for the device architecture
};

} // namespace bbfft

#endif // BBFFT_MODULE_FORMAT_HPP
