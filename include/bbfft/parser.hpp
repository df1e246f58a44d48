// FFT descriptor and device-info string parsers.
// Mirrors the reference interface include/bbfft/parser.hpp:24-34.
#ifndef BBFFT_PARSER_HPP
#define BBFFT_PARSER_HPP

#include "bbfft/configuration.hpp"
#include "bbfft/device_info.hpp"
#include "bbfft/export.hpp"

#include <string_view>

namespace bbfft {

/**
 * @brief Parse an FFT descriptor, e.g. "scfo8.256*32768" or "srfi128x128x128*32".
 *
 * Grammar: precision(s|d) domain(c|r) direction(f|b) placement(i|o)
 *          [M "."] N1 *2("x" Nd) ["*" K] ["i" stride] ["o" stride]
 * Throws std::runtime_error with a caret under the offending character when malformed.
 */
BBFFT_EXPORT configuration parse_fft_descriptor(std::string_view desc);
/// Parse "{max_wg, {sgs, ...}, local_mem, gpu|cpu}".
BBFFT_EXPORT device_info parse_device_info(std::string_view desc);

} // namespace bbfft

#endif // BBFFT_PARSER_HPP
