/*
 * FP64 DFT oracle for the double-batched FFT (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this code, and
 * only as a checker. The library never calls it.
 *
 * Definition (reference: docs/manual/plans.rst:12-19, include/bbfft/configuration.hpp:66-69):
 *   X(k) = sum_n x(n) exp(dir * 2 pi i * sum_d k_d n_d / N_d), dir = -1 forward, +1 backward,
 *   unnormalised, over the N modes of the column-major tensor M x N_1 x ... x N_d x K with
 *   element offset m s_0 + sum_d n_d s_d + k s_{d+1} (include/bbfft/configuration.hpp:162-179).
 * Real transforms (reference: src/base/generator/sbfft_gen.cpp:180-351):
 *   r2c stores bins n_1 = 0..floor(N_1/2) of mode 1;
 *   c2r reads those bins, ignores imag(X[0]) and, for even N_1, adds -imag(X[N_1/2]) to every
 *   output sample (the half-length pre-pass does not zero the Nyquist imaginary part,
 *   sbfft_gen.cpp:225-246 / f2fft_gen.cpp:340-366); multi-dimensional c2r first runs the complex
 *   backward transforms over modes 2..d and then the real transform over mode 1
 *   (src/common/algorithm/nd_fft.hpp:89-94).
 *
 * Method: brute-force O(N^2) DFT per mode with the exponent reduced exactly (n k mod N), a
 * twiddle table evaluated in long double, and long double accumulation.
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define MAXD 5

typedef struct {
    long double re, im;
} cld;

/* exp(dir 2 pi i r / n) for r = 0..n-1 in long double */
static cld *twiddles(size_t n, int dir) {
    cld *t = (cld *)malloc(sizeof(cld) * (n ? n : 1));
    const long double tau = 6.283185307179586476925286766559005768L;
    for (size_t r = 0; r < n; ++r) {
        long double a = tau * (long double)r / (long double)n;
        t[r].re = cosl(a);
        t[r].im = dir * sinl(a);
    }
    return t;
}

/* Dense complex work tensor with column-major shape s[0..nd-1]. */
typedef struct {
    unsigned nd;
    size_t s[MAXD];
    size_t st[MAXD];
    size_t size;
    cld *v;
} tensor;

static void tensor_init(tensor *t, unsigned nd, const size_t *s) {
    t->nd = nd;
    t->size = 1;
    for (unsigned i = 0; i < nd; ++i) {
        t->s[i] = s[i];
        t->st[i] = t->size;
        t->size *= s[i];
    }
    t->v = (cld *)calloc(t->size ? t->size : 1, sizeof(cld));
}

/* In-place DFT of length s[mode] along `mode` (brute force, long double). */
static void dft_mode(tensor *t, unsigned mode, int dir) {
    size_t n = t->s[mode], stride = t->st[mode];
    size_t inner = stride, outer = t->size / (n * stride);
    cld *tw = twiddles(n, dir);
    cld *col = (cld *)malloc(sizeof(cld) * (n ? n : 1));
    for (size_t o = 0; o < outer; ++o) {
        for (size_t i = 0; i < inner; ++i) {
            cld *base = t->v + o * n * stride + i;
            for (size_t j = 0; j < n; ++j) {
                col[j] = base[j * stride];
            }
            for (size_t k = 0; k < n; ++k) {
                long double re = 0, im = 0;
                size_t r = 0;
                for (size_t j = 0; j < n; ++j) {
                    re += col[j].re * tw[r].re - col[j].im * tw[r].im;
                    im += col[j].re * tw[r].im + col[j].im * tw[r].re;
                    r += k;
                    if (r >= n) {
                        r -= n;
                    }
                }
                base[k * stride].re = re;
                base[k * stride].im = im;
            }
        }
    }
    free(col);
    free(tw);
}

/* Unravel linear index of a dense shape into a multi-index. */
static void unravel(size_t lin, unsigned nd, const size_t *s, size_t *idx) {
    for (unsigned i = 0; i < nd; ++i) {
        idx[i] = lin % s[i];
        lin /= s[i];
    }
}

static size_t offset(unsigned nd, const size_t *idx, const size_t *stride) {
    size_t o = 0;
    for (unsigned i = 0; i < nd; ++i) {
        o += idx[i] * stride[i];
    }
    return o;
}

/*
 * type: 0 c2c, 1 r2c, 2 c2r. Complex data are interleaved (re, im) doubles; offsets follow the
 * strides in units of the element (real or complex). Returns 0 on success.
 */
int oracle_fft(unsigned dim, const size_t *shape, const size_t *istride, const size_t *ostride,
               int type, int dir, const double *in, double *out) {
    if (dim < 1 || dim > 3 || type < 0 || type > 2 || (dir != -1 && dir != 1)) {
        return 1;
    }
    unsigned nd = dim + 2;
    size_t full[MAXD], half[MAXD], idx[MAXD];
    for (unsigned i = 0; i < nd; ++i) {
        full[i] = half[i] = shape[i];
    }
    half[1] = shape[1] / 2 + 1;
    size_t N1 = shape[1];
    tensor t;
    if (type == 0) {
        tensor_init(&t, nd, full);
        for (size_t l = 0; l < t.size; ++l) {
            unravel(l, nd, full, idx);
            size_t o = offset(nd, idx, istride);
            t.v[l].re = in[2 * o];
            t.v[l].im = in[2 * o + 1];
        }
        for (unsigned d = 1; d <= dim; ++d) {
            dft_mode(&t, d, dir);
        }
        for (size_t l = 0; l < t.size; ++l) {
            unravel(l, nd, full, idx);
            size_t o = offset(nd, idx, ostride);
            out[2 * o] = (double)t.v[l].re;
            out[2 * o + 1] = (double)t.v[l].im;
        }
    } else if (type == 1) {
        tensor_init(&t, nd, full);
        for (size_t l = 0; l < t.size; ++l) {
            unravel(l, nd, full, idx);
            t.v[l].re = in[offset(nd, idx, istride)];
            t.v[l].im = 0;
        }
        for (unsigned d = 1; d <= dim; ++d) {
            dft_mode(&t, d, dir);
        }
        size_t hsize = 1;
        for (unsigned i = 0; i < nd; ++i) {
            hsize *= half[i];
        }
        for (size_t l = 0; l < hsize; ++l) {
            unravel(l, nd, half, idx);
            size_t o = offset(nd, idx, ostride);
            cld v = t.v[offset(nd, idx, t.st)];
            out[2 * o] = (double)v.re;
            out[2 * o + 1] = (double)v.im;
        }
    } else {
        /* complex backward transforms over modes 2..d on the half spectrum */
        tensor h;
        tensor_init(&h, nd, half);
        for (size_t l = 0; l < h.size; ++l) {
            unravel(l, nd, half, idx);
            size_t o = offset(nd, idx, istride);
            h.v[l].re = in[2 * o];
            h.v[l].im = in[2 * o + 1];
        }
        for (unsigned d = 2; d <= dim; ++d) {
            dft_mode(&h, d, dir);
        }
        /* real transform over mode 1 with the reference's DC / Nyquist handling */
        const long double tau = 6.283185307179586476925286766559005768L;
        size_t H = half[1];
        size_t inner = h.st[1], outer = h.size / (H * inner);
        tensor_init(&t, nd, full);
        for (size_t o = 0; o < outer; ++o) {
            for (size_t i = 0; i < inner; ++i) {
                const cld *X = h.v + o * H * inner + i;
                cld *x = t.v + o * N1 * inner + i;
                for (size_t n = 0; n < N1; ++n) {
                    long double acc = X[0].re;
                    for (size_t k = 1; k < H; ++k) {
                        long double a = tau * (long double)((n * k) % N1) / (long double)N1;
                        long double c = cosl(a), s = dir * sinl(a);
                        cld v = X[k * inner];
                        if (N1 % 2 == 0 && k == N1 / 2) {
                            acc += v.re * c - v.im; /* c = (-1)^n */
                        } else {
                            acc += 2 * (v.re * c - v.im * s);
                        }
                    }
                    x[n * inner].re = acc;
                    x[n * inner].im = 0;
                }
            }
        }
        for (size_t l = 0; l < t.size; ++l) {
            unravel(l, nd, full, idx);
            out[offset(nd, idx, ostride)] = (double)t.v[l].re;
        }
        free(h.v);
    }
    free(t.v);
    return 0;
}
