// Errors raised by the HIP back end.
// Plays the role of the reference's include/bbfft/cl/error.hpp:20-69 / ze/error.hpp.
#ifndef BBFFT_HIP_ERROR_HPP
#define BBFFT_HIP_ERROR_HPP

#include "bbfft/export.hpp"

#include <hip/hip_runtime_api.h>
#include <hip/hiprtc.h>

#include <stdexcept>
#include <string>

namespace bbfft::hip {

/// HIP runtime or hiprtc failure; what() carries the call, the status and a compiler log.
class BBFFT_EXPORT error : public std::runtime_error {
  public:
    error(std::string const &what, int status);
    int status() const noexcept { return status_; }

  private:
    int status_;
};

BBFFT_EXPORT void check(hipError_t status, char const *what);
BBFFT_EXPORT void check(hiprtcResult status, char const *what, std::string const &log = {});

} // namespace bbfft::hip

#define BBFFT_HIP_CHECK(X) ::bbfft::hip::check((X), #X)

#endif // BBFFT_HIP_ERROR_HPP
