#include "math.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace bbfft {

int ipow(int base, unsigned exp) {
    int r = 1;
    for (unsigned i = 0; i < exp; ++i) {
        r *= base;
    }
    return r;
}

unsigned iroot(unsigned radicand, unsigned index) {
    // Binary search for the largest r with r^index <= radicand (overflow-safe).
    auto fits = [&](unsigned long long r) {
        unsigned long long p = 1;
        for (unsigned i = 0; i < index; ++i) {
            p *= r;
            if (p > radicand) {
                return false;
            }
        }
        return true;
    };
    unsigned long long lo = 0, hi = static_cast<unsigned long long>(radicand) + 1;
    while (hi - lo > 1) {
        auto mid = (lo + hi) / 2;
        if (fits(mid)) {
            lo = mid;
        } else {
            hi = mid;
        }
    }
    return static_cast<unsigned>(lo);
}

bool is_prime(unsigned n) {
    if (n < 2) {
        return false;
    }
    for (unsigned d = 2; d * d <= n; ++d) {
        if (n % d == 0) {
            return false;
        }
    }
    return true;
}

std::size_t max_power_of_2_less_equal(std::size_t x) {
    std::size_t p = 1;
    while (p <= x / 2) {
        p *= 2;
    }
    return p;
}

std::size_t min_power_of_2_greater_equal(std::size_t x) {
    std::size_t p = 1;
    while (p < x) {
        p *= 2;
    }
    return p;
}

std::vector<int> trial_division(int n) {
    std::vector<int> f;
    for (int d = 2; n > 1;) {
        if (n % d == 0) {
            f.push_back(d);
            n /= d;
        } else {
            ++d;
        }
    }
    return f;
}

namespace {
// Best split of n into `index` factors: minimises the sum of squared distances of the factors to
// `target`. Candidates for the first factor run downwards from floor(n^(1/index)); the error of a
// candidate is the best error of the remaining split plus its own squared distance, and the first
// strictly smaller error wins.
double best_split(unsigned n, unsigned index, double target, std::vector<unsigned> &out) {
    if (index == 1) {
        out.assign(1, n);
        double d = target - n;
        return d * d;
    }
    double best_err = std::numeric_limits<double>::max();
    std::vector<unsigned> sub;
    for (unsigned f = iroot(n, index); f > 0; --f) {
        if (n % f != 0) {
            continue;
        }
        double err = best_split(n / f, index - 1, target, sub);
        double d = target - f;
        err += d * d;
        if (err < best_err) {
            best_err = err;
            out.assign(1, f);
            out.insert(out.end(), sub.begin(), sub.end());
        }
    }
    return best_err;
}
} // namespace

std::vector<unsigned> factor(unsigned n, unsigned index) {
    if (index == 0) {
        return {};
    }
    if (n == 0) {
        return std::vector<unsigned>(index, 0);
    }
    std::vector<unsigned> f;
    best_split(n, index, std::pow(static_cast<double>(n), 1.0 / index), f);
    std::sort(f.begin(), f.end());
    return f;
}

} // namespace bbfft
