// Device queries (reference pattern: src/cl/device.cpp:12-40, src/ze/device.cpp:12-43).
#include "bbfft/hip/device.hpp"
#include "bbfft/hip/error.hpp"

#include <hip/hip_runtime_api.h>

#include <functional>
#include <string>

namespace bbfft::hip {

namespace {
hipDeviceProp_t props(int device) {
    hipDeviceProp_t p;
    BBFFT_HIP_CHECK(hipGetDeviceProperties(&p, device));
    return p;
}
std::string arch_of(hipDeviceProp_t const &p) {
    std::string a = p.gcnArchName;
    auto colon = a.find(':');
    return colon == std::string::npos ? a : a.substr(0, colon);
}
} // namespace

device_info get_device_info(int device) {
    auto p = props(device);
    device_info info;
    info.max_work_group_size = static_cast<std::size_t>(p.maxThreadsPerBlock);
    info.subgroup_sizes = {static_cast<std::size_t>(p.warpSize)};
    info.local_memory_size = p.sharedMemPerBlock;
    info.type = device_type::gpu;
    return info;
}

std::uint64_t get_device_id(int device) {
    auto p = props(device);
    std::uint64_t id = (static_cast<std::uint64_t>(p.pciDomainID) << 40) ^
                       (static_cast<std::uint64_t>(p.pciBusID) << 16) ^
                       static_cast<std::uint64_t>(p.pciDeviceID);
    id ^= static_cast<std::uint64_t>(std::hash<std::string>{}(arch_of(p))) << 1;
    return id;
}

std::string get_device_arch(int device) { return arch_of(props(device)); }

} // namespace bbfft::hip
