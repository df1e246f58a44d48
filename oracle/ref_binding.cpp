// C binding over reference sources that compile on their own (TEST INFRASTRUCTURE ONLY).
// Built by oracle/ref.mk from the reference tree into oracle/_ref/libref.so; used by the CPU
// tests to check the library's and the oracle port's integer logic against the reference's:
//   factor / iroot / ipow / is_prime  (src/base/prime_factorization.cpp, src/base/math.cpp)
//   scrambler / unscrambler           (src/base/scrambler.hpp, header only)
#include "math.hpp"
#include "prime_factorization.hpp"
#include "scrambler.hpp"

#include <vector>

extern "C" {

void ref_factor(unsigned n, unsigned index, unsigned *out) {
    auto f = bbfft::factor(n, index);
    for (unsigned i = 0; i < f.size(); ++i) {
        out[i] = f[i];
    }
}

unsigned ref_iroot(unsigned radicand, unsigned index) { return bbfft::iroot(radicand, index); }
int ref_ipow(int base, unsigned exp) { return bbfft::ipow(base, exp); }
int ref_is_prime(unsigned n) { return bbfft::is_prime(n) ? 1 : 0; }

int ref_scramble(const int *fac, int nf, int index, int in0toN) {
    bbfft::scrambler<int> s(std::vector<int>(fac, fac + nf));
    s.in0toN(in0toN != 0);
    return s(index);
}

int ref_unscramble(const int *fac, int nf, int index, int in0toN) {
    bbfft::unscrambler<int> s(std::vector<int>(fac, fac + nf));
    s.in0toN(in0toN != 0);
    return s(index);
}

int ref_trial_division(int n, int *out) {
    auto f = bbfft::trial_division(n);
    for (std::size_t i = 0; i < f.size(); ++i) {
        out[i] = f[i];
    }
    return static_cast<int>(f.size());
}
}
