// On-disk store of ahead-of-time compiled code objects (written by bbfft-aot-generate).
//
// Layout (little endian): "BBFFTAOT", u32 version (1), u32 entry count; per entry:
// u32 arch length, arch bytes, u32 name count, (u32 length, name bytes) per name,
// u64 code-object size, code-object bytes.
#ifndef BBFFT_AMD_AOT_ARCHIVE_HPP
#define BBFFT_AMD_AOT_ARCHIVE_HPP

#include <cstdint>
#include <string>
#include <vector>

namespace bbfft::hip {

struct archive_entry {
    std::string arch;
    std::vector<std::string> kernel_names;
    std::vector<std::uint8_t> code;
};

void write_archive(std::string const &path, std::vector<archive_entry> const &entries);
std::vector<archive_entry> read_archive(std::string const &path);
/// Parse an archive held in memory (`what` names it in error messages).
std::vector<archive_entry> parse_archive(std::uint8_t const *data, std::size_t size,
                                         std::string const &what);

} // namespace bbfft::hip

#endif // BBFFT_AMD_AOT_ARCHIVE_HPP
