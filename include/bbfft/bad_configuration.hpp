// Exception for unsupported or faulty configurations.
// Mirrors the reference interface include/bbfft/bad_configuration.hpp:17-40.
#ifndef BBFFT_BAD_CONFIGURATION_HPP
#define BBFFT_BAD_CONFIGURATION_HPP

#include "bbfft/export.hpp"

#include <exception>
#include <string>

namespace bbfft {

class BBFFT_EXPORT bad_configuration : public std::exception {
  public:
    bad_configuration(std::string what);
    bad_configuration(char const *what);
    char const *what() const noexcept override;

  private:
    std::string what_;
};

} // namespace bbfft

#endif // BBFFT_BAD_CONFIGURATION_HPP
