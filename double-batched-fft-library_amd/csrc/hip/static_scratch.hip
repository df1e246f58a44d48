// Device-resident scratch array of the allocation-free execute entry point
// (bbfft_plan_execute_static_scratch). Multi-dimensional plans whose input is larger than their
// output (out-of-place c2r) keep the intermediate tensor between passes in a scratch buffer;
// the plan allocates its own at creation, while the allocation-free entry point uses this
// file-scope array instead, so it never calls hipMalloc/hipFree.
//
// Each GPU holds its own copy of the array (at its own device address) from first use until
// the process ends. `scratch` below names a host-side stand-in in host code, not the device
// address: the address is obtained per device with hipGetSymbolAddress and only that pointer
// is handed to kernels.
#include "static_scratch.hpp"

#include "bbfft/bad_configuration.hpp"
#include "bbfft/hip/error.hpp"

#include <hip/hip_runtime.h>

#include <mutex>

namespace bbfft::hip_detail {

namespace {

__device__ __align__(256) unsigned char scratch[static_scratch_bytes];

constexpr int max_devices = 64;
std::mutex g_mutex;
void *g_ptr[max_devices] = {};

} // namespace

void *static_scratch(int device) {
    if (device < 0 || device >= max_devices) {
        throw bad_configuration("Device ordinal out of range for the static scratch array");
    }
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_ptr[device]) {
        int current = 0;
        BBFFT_HIP_CHECK(hipGetDevice(&current));
        if (current != device) {
            BBFFT_HIP_CHECK(hipSetDevice(device));
        }
        void *ptr = nullptr;
        hipError_t const err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (current != device) {
            BBFFT_HIP_CHECK(hipSetDevice(current));
        }
        BBFFT_HIP_CHECK(err);
        if (!ptr) {
            throw hip::error("hipGetSymbolAddress returned no address for the scratch array",
                             hipErrorInvalidSymbol);
        }
        g_ptr[device] = ptr;
    }
    return g_ptr[device];
}

} // namespace bbfft::hip_detail
