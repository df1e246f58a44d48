/*
 * CPU port of the reference's 1-D c2c algorithm (TEST INFRASTRUCTURE / CPU BASELINE ONLY).
 *
 * Only tests/ and bench.py's cpu_baseline leg load this code. It restates, in plain C on host
 * cores, what the reference runs for a 1-D c2c plan on its CPU device_info
 * {8192, {4, 8, 16, 32, 64}, 32768, cpu} (reference test/parser.cpp:27):
 *
 *   dispatch      src/common/algorithm_1d.hpp:19-36 (register-space rule -> sbfft or f2fft)
 *   tiling        src/base/generator/small_batch_fft.cpp:15-58, factor2_slm_fft.cpp:16-86
 *   factor        src/base/prime_factorization.cpp:13-77, src/base/math.cpp:8-71
 *   f2fft stages  src/base/generator/f2fft_gen.cpp:95-176 (stage f = L-1..0, stride-J1 loads,
 *                 in-register DFT, twiddle from the table, unscrambled store to SLM) and the
 *                 c2c post-pass f2fft_gen.cpp:203-226
 *   sbfft         src/base/generator/sbfft_gen.cpp:64-127 (whole column in registers)
 *   in-register   src/base/mixed_radix_fft.cpp:164-197 (inplace<ESum>) with the pair
 *                 optimisation of mixed_radix_fft.cpp:56-118
 *   twiddles      src/common/algorithm/factor2_slm_fft.hpp:49-90 (double, cast to T)
 *   constants     src/base/root_of_unity.cpp:144-189 (power_of_w: reduced fraction; correctly
 *                 rounded values for n <= 64, cos/sin in double otherwise)
 *   scrambler     src/base/scrambler.hpp:16-101
 *
 * Work-items of the M block are the innermost loop (the lanes a CPU OpenCL compiler vectorises);
 * work-groups are distributed over pthreads.
 */
#include <math.h>
#include <pthread.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define MAXF 8
#define MAXLANES 16
#define MAXPAIRS 1024

/* ---- integer helpers (src/base/math.cpp, src/base/prime_factorization.cpp) -------------- */

static unsigned iroot_u(unsigned radicand, unsigned index) {
    unsigned long long lo = 0, hi = (unsigned long long)radicand + 1;
    while (hi - lo > 1) {
        unsigned long long mid = (lo + hi) / 2, p = 1;
        int fits = 1;
        for (unsigned i = 0; i < index; ++i) {
            p *= mid;
            if (p > radicand) {
                fits = 0;
                break;
            }
        }
        if (fits) {
            lo = mid;
        } else {
            hi = mid;
        }
    }
    return (unsigned)lo;
}

static int is_prime_u(unsigned n) {
    if (n < 2) {
        return 0;
    }
    for (unsigned d = 2; d * d <= n; ++d) {
        if (n % d == 0) {
            return 0;
        }
    }
    return 1;
}

static size_t pow2_ge(size_t x) {
    size_t p = 1;
    while (p < x) {
        p *= 2;
    }
    return p;
}

static int trial_division_i(int n, int *f) {
    int nf = 0;
    for (int d = 2; n > 1;) {
        if (n % d == 0) {
            f[nf++] = d;
            n /= d;
        } else {
            ++d;
        }
    }
    return nf;
}

/* factor(n, index): best split by squared distance to n^(1/index), first strictly best wins */
static double best_split(unsigned n, unsigned index, double target, unsigned *out) {
    if (index == 1) {
        out[0] = n;
        double d = target - n;
        return d * d;
    }
    double best = 1e308;
    unsigned sub[MAXF];
    for (unsigned f = iroot_u(n, index); f > 0; --f) {
        if (n % f) {
            continue;
        }
        double err = best_split(n / f, index - 1, target, sub);
        double d = target - f;
        err += d * d;
        if (err < best) {
            best = err;
            out[0] = f;
            memcpy(out + 1, sub, sizeof(unsigned) * (index - 1));
        }
    }
    return best;
}

static int cmp_u(const void *a, const void *b) {
    unsigned x = *(const unsigned *)a, y = *(const unsigned *)b;
    return (x > y) - (x < y);
}

void cpu_factor(unsigned n, unsigned index, unsigned *out) {
    if (index == 0) {
        return;
    }
    if (n == 0) {
        memset(out, 0, sizeof(unsigned) * index);
        return;
    }
    best_split(n, index, pow((double)n, 1.0 / index), out);
    qsort(out, index, sizeof(unsigned), cmp_u);
}

/* ---- scrambler / unscrambler (src/base/scrambler.hpp) ------------------------------------ */

int cpu_scramble(const int *fac, int nf, int index, int in0toN) {
    if (nf == 1 && in0toN) {
        return index;
    }
    int r = 0, n = 1;
    for (int i = 0; i < nf; ++i) {
        r = r * fac[i] + index % fac[i];
        index /= fac[i];
        n *= fac[i];
    }
    return in0toN ? r : r + index * n;
}

int cpu_unscramble(const int *fac, int nf, int index, int in0toN) {
    if (nf == 1 && in0toN) {
        return index;
    }
    int r = 0, n = 1;
    for (int i = nf - 1; i >= 0; --i) {
        r = r * fac[i] + index % fac[i];
        index /= fac[i];
        n *= fac[i];
    }
    return in0toN ? r : r + index * n;
}

/* ---- power_of_w (src/base/root_of_unity.cpp:144-189) -------------------------------------- */

static long gcd_l(long a, long b) {
    if (a < 0) {
        a = -a;
    }
    if (b < 0) {
        b = -b;
    }
    while (b) {
        long t = a % b;
        a = b;
        b = t;
    }
    return a;
}

/* exp(2 pi i k / n): the fraction is reduced and mapped to |k| <= n/2; n <= 64 uses correctly
 * rounded values (evaluated in long double here), larger n uses cos/sin in double. */
void cpu_power_of_w(long k, long n, double *re, double *im) {
    k %= n;
    long g = gcd_l(k, n);
    if (g == 0) {
        *re = 1.0;
        *im = 0.0;
        return;
    }
    k /= g;
    n /= g;
    if (2 * k > n) {
        k -= n;
    } else if (2 * k < -n) {
        k += n;
    }
    if (n <= 64) {
        /* correctly rounded: exact on the axes, long double otherwise (|k| <= n/2 here) */
        long kp = k < 0 ? -k : k;
        if (4 * kp == n) {
            *re = 0.0;
            *im = 1.0;
        } else if (2 * kp == n) {
            *re = -1.0;
            *im = 0.0;
        } else {
            const long double tau = 6.283185307179586476925286766559005768L;
            long double a = tau * (long double)kp / (long double)n;
            *re = (double)cosl(a);
            *im = (double)sinl(a);
        }
        if (k < 0) {
            *im = -*im;
        }
        return;
    }
    double a = (6.28318530717958647693 / n) * k;
    *re = cos(a);
    *im = sin(a);
}

/* ---- precision-generic part --------------------------------------------------------------- */

#define T float
#define SFX _f32
#include "bbfft_cpu_impl.h"
#undef T
#undef SFX
#define T double
#define SFX _f64
#include "bbfft_cpu_impl.h"
#undef T
#undef SFX
