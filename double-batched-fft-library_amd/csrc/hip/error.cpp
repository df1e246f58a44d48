#include "bbfft/hip/error.hpp"

#include <sstream>

namespace bbfft::hip {

error::error(std::string const &what, int status) : std::runtime_error(what), status_(status) {}

void check(hipError_t status, char const *what) {
    if (status != hipSuccess) {
        std::ostringstream oss;
        oss << what << " failed: " << hipGetErrorName(status) << " (" << static_cast<int>(status)
            << "): " << hipGetErrorString(status);
        throw error(oss.str(), static_cast<int>(status));
    }
}

void check(hiprtcResult status, char const *what, std::string const &log) {
    if (status != HIPRTC_SUCCESS) {
        std::ostringstream oss;
        oss << what << " failed: " << hiprtcGetErrorString(status) << " ("
            << static_cast<int>(status) << ")";
        if (!log.empty()) {
            oss << "\nCompiler log:\n" << log;
        }
        throw error(oss.str(), static_cast<int>(status));
    }
}

} // namespace bbfft::hip
