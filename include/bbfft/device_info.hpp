// Properties of the target device used to choose kernel tilings.
// Mirrors the reference interface include/bbfft/device_info.hpp:18-59.
#ifndef BBFFT_DEVICE_INFO_HPP
#define BBFFT_DEVICE_INFO_HPP

#include "bbfft/export.hpp"

#include <cstddef>
#include <iosfwd>
#include <string>
#include <vector>

namespace bbfft {

enum class BBFFT_EXPORT device_type {
    gpu,   ///< GPU
    cpu,   ///< CPU
    custom ///< other
};

/**
 * @brief Tiling inputs of a device.
 *
 * For gfx950 (MI355X): {1024, {64}, 163840, gpu}, i.e. 1024 work-items per work-group, wave64,
 * 160 KiB of LDS per compute unit.
 */
struct BBFFT_EXPORT device_info {
    std::size_t max_work_group_size = 0;     ///< work-items per work-group
    std::vector<std::size_t> subgroup_sizes; ///< supported sub-group (wavefront) sizes
    std::size_t local_memory_size = 0;       ///< bytes of local (LDS) memory
    device_type type = device_type::gpu;     ///< device type

    std::size_t min_subgroup_size() const;
    std::size_t max_subgroup_size() const;
    /// Bytes of vector register file available to one sub-group.
    std::size_t register_space_min() const;
    std::size_t register_space_max() const;
    std::string to_string() const;
    bool operator==(device_info const &other) const;
    bool operator!=(device_info const &other) const;
};

BBFFT_EXPORT std::ostream &operator<<(std::ostream &os, device_type type);
BBFFT_EXPORT std::ostream &operator<<(std::ostream &os, device_info const &info);

} // namespace bbfft

#endif // BBFFT_DEVICE_INFO_HPP
