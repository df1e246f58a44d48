// device_info helpers.
// Reference behaviour: src/base/device_info.cpp:13-95. The GPU register model is replaced by the
// CDNA one: a wave64 owns up to 512 VGPRs of 4 bytes per lane.
#include "bbfft/device_info.hpp"

#include <algorithm>
#include <ostream>
#include <sstream>
#include <stdexcept>

namespace bbfft {

std::size_t device_info::min_subgroup_size() const {
    return subgroup_sizes.empty() ? 8 : *std::min_element(subgroup_sizes.begin(),
                                                          subgroup_sizes.end());
}

std::size_t device_info::max_subgroup_size() const {
    return subgroup_sizes.empty() ? 8 : *std::max_element(subgroup_sizes.begin(),
                                                          subgroup_sizes.end());
}

std::size_t device_info::register_space_min() const {
    switch (type) {
    case device_type::cpu:
        return 32 * 64; // 32 zmm registers of 64 bytes (AVX-512)
    case device_type::gpu:
        if (min_subgroup_size() >= 64) {
            return 64 * 4 * 256; // CDNA wave64: 256 arch VGPRs x 4 bytes x 64 lanes
        }
        // Sub-group-32/16 GPUs: 128 registers of 32 bytes scaled by sub-group width.
        return std::max<std::size_t>(1, min_subgroup_size() / 8) * 32 * 128;
    case device_type::custom:
        throw std::runtime_error("register_space unknown for custom device");
    }
    return 0;
}

std::size_t device_info::register_space_max() const { return register_space_min(); }

std::string device_info::to_string() const {
    std::ostringstream oss;
    oss << *this;
    return oss.str();
}

bool device_info::operator==(device_info const &other) const {
    return max_work_group_size == other.max_work_group_size &&
           subgroup_sizes == other.subgroup_sizes &&
           local_memory_size == other.local_memory_size && type == other.type;
}
bool device_info::operator!=(device_info const &other) const { return !(*this == other); }

std::ostream &operator<<(std::ostream &os, device_type type) {
    switch (type) {
    case device_type::gpu:
        return os << "gpu";
    case device_type::cpu:
        return os << "cpu";
    default:
        return os << "custom";
    }
}

std::ostream &operator<<(std::ostream &os, device_info const &info) {
    os << '{' << info.max_work_group_size << ", {";
    for (std::size_t i = 0; i < info.subgroup_sizes.size(); ++i) {
        os << (i ? ", " : "") << info.subgroup_sizes[i];
    }
    return os << "}, " << info.local_memory_size << ", " << info.type << '}';
}

} // namespace bbfft
