// Device queries for the HIP back end.
// Plays the role of the reference's include/bbfft/{cl,ze,sycl}/device.hpp.
#ifndef BBFFT_HIP_DEVICE_HPP
#define BBFFT_HIP_DEVICE_HPP

#include "bbfft/device_info.hpp"
#include "bbfft/export.hpp"

#include <cstdint>
#include <string>

namespace bbfft::hip {

/// Tiling inputs of a HIP device (from hipDeviceProp_t).
BBFFT_EXPORT device_info get_device_info(int device);
/// Stable id of a device: PCI domain/bus/device and the architecture name.
BBFFT_EXPORT std::uint64_t get_device_id(int device);
/// Architecture name, e.g. "gfx950".
BBFFT_EXPORT std::string get_device_arch(int device);

} // namespace bbfft::hip

#endif // BBFFT_HIP_DEVICE_HPP
