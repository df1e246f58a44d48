// Persisted autotuning table: per (precision, transform type, N, M) the kernel parameters the
// autotuner measured fastest on gfx950. The type may carry the input storage of half-input
// kernels: c2c-f16, c2c-bf16, c2c-f16-mfma, c2c-bf16-mfma; c2c-ragged entries hold the
// variants of the persistent ragged kernel. Built into the library from tuning/gfx950.inc;
// BBFFT_TUNING_FILE replaces it with a table file of the same format and BBFFT_TUNING=off
// disables it (heuristics only).
#include "tuning_table.hpp"

#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>

namespace bbfft::gen {

namespace {

char const builtin_table[] =
#include "../../tuning/gfx950.inc"
    ;

using key_t = std::tuple<int, int, std::size_t, std::size_t, unsigned>;

/// Input-storage bits of a kernel variant (and the ragged-variant bit) that select a separate
/// table entry.
constexpr unsigned storage_bits = fft1d_input_f16 | fft1d_input_bf16 | fft1d_mfma | tuning_ragged;

std::map<key_t, fft1d_tuning> parse(std::string const &text) {
    std::map<key_t, fft1d_tuning> table;
    std::istringstream lines(text);
    std::string line;
    while (std::getline(lines, line)) {
        auto hash = line.find('#');
        if (hash != std::string::npos) {
            line.resize(hash);
        }
        auto colon = line.find(':');
        if (colon == std::string::npos) {
            continue;
        }
        std::istringstream lhs(line.substr(0, colon)), rhs(line.substr(colon + 1));
        std::string prec, type, fac;
        std::size_t N = 0, M = 0;
        fft1d_tuning t;
        if (!(lhs >> prec >> type >> N >> M) ||
            !(rhs >> fac >> t.Nt >> t.Mb >> t.Kb >> t.lds_pad >> t.flags)) {
            continue;
        }
        int p = prec == "f32" ? 4 : prec == "f64" ? 8 : 0;
        unsigned storage = 0;
        for (auto [suffix, bit] : {std::pair<char const *, unsigned>{"-f16", fft1d_input_f16},
                                   {"-bf16", fft1d_input_bf16},
                                   {"-mfma", fft1d_mfma},
                                   {"-ragged", tuning_ragged}}) {
            auto pos = type.find(suffix);
            if (pos != std::string::npos) {
                storage |= bit;
                type.erase(pos, std::string(suffix).size());
            }
        }
        int ty = type == "c2c" ? 0 : type == "r2c" ? 1 : type == "c2r" ? 2 : -1;
        if (!p || ty < 0) {
            continue;
        }
        std::istringstream fs(fac);
        std::string f;
        while (std::getline(fs, f, 'x')) {
            t.factorization.push_back(std::stoi(f));
        }
        unsigned waves = 0, groups = 0;
        if (rhs >> waves) { // optional occupancy target
            t.waves_per_eu = waves;
            if (rhs >> groups) { // optional persistent grid
                t.groups_per_cu = groups;
            }
        }
        t.no_pad = t.lds_pad == 0;
        t.has_flags = true;
        table[key_t{p, ty, N, M, storage}] = t;
    }
    return table;
}

std::map<key_t, fft1d_tuning> const &table() {
    static std::map<key_t, fft1d_tuning> tab;
    static std::once_flag once;
    std::call_once(once, [] {
        char const *off = std::getenv("BBFFT_TUNING");
        if (off && std::string(off) == "off") {
            return;
        }
        char const *file = std::getenv("BBFFT_TUNING_FILE");
        if (file && *file) {
            std::ifstream f(file);
            std::stringstream ss;
            ss << f.rdbuf();
            tab = parse(ss.str());
        } else {
            tab = parse(builtin_table);
        }
    });
    return tab;
}

} // namespace

bool lookup_tuning(precision fp, transform_type type, std::size_t N, std::size_t M,
                   unsigned variant, fft1d_tuning &out) {
    auto const &t = table();
    auto it = t.find(
        key_t{static_cast<int>(fp), static_cast<int>(type), N, M, variant & storage_bits});
    if (it == t.end()) {
        return false;
    }
    out = it->second;
    return true;
}

std::size_t tuning_table_size() { return table().size(); }

} // namespace bbfft::gen
