// Offline generation of every kernel a set of configurations needs.
// Reference behaviour: src/base/generator.cpp:16-25 (there via a recording back end).
#include "bbfft/generator.hpp"
#include "bbfft/detail/generator_impl.hpp"
#include "decompose.hpp"

#include <ostream>
#include <set>

namespace bbfft {

std::vector<std::string> generate_fft_kernels(std::ostream &os,
                                              std::vector<configuration> const &cfgs,
                                              device_info const &info) {
    std::vector<std::string> names;
    std::set<std::string> seen;
    for (auto const &cfg : cfgs) {
        for (auto const &pass : decompose(cfg).passes) {
            auto kc = configure_fft1d(pass, info);
            auto id = kc.identifier();
            if (seen.insert(id).second) {
                if (pass.callbacks) {
                    os << std::string(pass.callbacks.data, pass.callbacks.length) << '\n';
                }
                generate_fft1d(os, kc);
                names.push_back(id);
            }
        }
    }
    return names;
}

} // namespace bbfft
