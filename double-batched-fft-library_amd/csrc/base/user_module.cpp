#include "bbfft/user_module.hpp"

namespace bbfft {

user_module::operator bool() const noexcept {
    return data != nullptr && length > 0 && (load_function != nullptr || store_function != nullptr);
}

} // namespace bbfft
