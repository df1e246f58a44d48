// Offline kernel generation (feeds the ahead-of-time store).
// Mirrors the reference interface include/bbfft/generator.hpp:26-28; the output is HIP source.
#ifndef BBFFT_GENERATOR_HPP
#define BBFFT_GENERATOR_HPP

#include "bbfft/configuration.hpp"
#include "bbfft/device_info.hpp"
#include "bbfft/export.hpp"

#include <iosfwd>
#include <string>
#include <vector>

namespace bbfft {

/// Write the HIP source of every kernel the plans of cfgs need; returns the kernel names.
BBFFT_EXPORT std::vector<std::string> generate_fft_kernels(std::ostream &os,
                                                           std::vector<configuration> const &cfgs,
                                                           device_info const &info);

} // namespace bbfft

#endif // BBFFT_GENERATOR_HPP
