#include "roots.hpp"

#include <quadmath.h>

namespace bbfft::gen {

namespace {

struct qcomplex {
    __float128 re, im;
};

// exp(2 pi i k / n) in quad precision with exact octant reduction.
qcomplex root_q(long k, long n) {
    static __float128 const half_pi = acosq(__float128(0));
    if (n < 0) {
        n = -n;
        k = -k;
    }
    long r = k % n;
    if (r < 0) {
        r += n;
    }
    // angle = 2 pi r / n = (pi/2) (q + f), q quadrant, f = rem / n in [0, 1)
    long const four_r = 4 * r;
    long const q = four_r / n;
    long rem = four_r - q * n; // angle within quadrant = (pi/2) rem / n
    __float128 c, s;
    if (rem == 0) {
        c = 1;
        s = 0;
    } else if (2 * rem == n) {
        c = s = sqrtq(__float128(0.5));
    } else if (2 * rem < n) {
        __float128 a = half_pi * (__float128(rem) / n);
        c = cosq(a);
        s = sinq(a);
    } else {
        __float128 a = half_pi * (__float128(n - rem) / n);
        c = sinq(a);
        s = cosq(a);
    }
    switch (q & 3) {
    case 0:
        return {c, s};
    case 1:
        return {-s, c};
    case 2:
        return {-c, -s};
    default:
        return {s, -c};
    }
}

} // namespace

std::complex<double> root_of_unity(long k, long n) {
    auto w = root_q(k, n);
    return {static_cast<double>(w.re), static_cast<double>(w.im)};
}

std::complex<float> root_of_unity_f(long k, long n) {
    auto w = root_q(k, n);
    return {static_cast<float>(w.re), static_cast<float>(w.im)};
}

} // namespace bbfft::gen
