/*
 * Precision-generic body of the CPU port (included twice by bbfft_cpu.c with T = float/double).
 * TEST INFRASTRUCTURE / CPU BASELINE ONLY. Reference files are cited in bbfft_cpu.c.
 */
#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)

typedef struct {
    int N, M, K, dir;
    size_t is1, is2, os1, os2;
    int sbfft;           /* 1: one work-item per column (small_batch_fft) */
    int fac[MAXF], L;    /* f2fft factorisation (stage order f = L-1 .. 0) */
    int Mb, Kb;          /* work-group tile */
    T *twiddle;          /* f2fft twiddle table, interleaved (re, im) */
} FN(cpu_plan);

/* lanes of one complex value: index [lane] */
typedef struct {
    T re[MAXLANES];
    T im[MAXLANES];
} FN(cv);

/* Pair-optimised DFT sum for output kf (mixed_radix_fft.cpp:56-118); x[jf] are lane vectors.
 * pair_v1 / pair_v2 cache products of conjugate pairs across kf (reuse for -kf). */
typedef struct {
    int n; /* number of cached pairs */
    long key[MAXPAIRS][4];
    FN(cv) v1[MAXPAIRS];
    FN(cv) v2[MAXPAIRS];
    /* power_of_w(i, pwN) for i = 0..pwN-1: every root the DFTs of one plan use has a
     * denominator dividing pwN and power_of_w depends only on the reduced fraction, so this
     * table returns exactly power_of_w(k, n) (the reference bakes these values into the kernel
     * source as literals; the port looks them up instead of re-evaluating them per element) */
    const double *pw;
    long pwN;
} FN(pair_cache);

static void FN(power_of_w_t)(FN(pair_cache) const *pc, long k, long n, double *re, double *im) {
    if (pc->pw && n > 0 && pc->pwN % n == 0) {
        long r = k % n;
        if (r < 0) {
            r += n;
        }
        long i = r * (pc->pwN / n);
        *re = pc->pw[2 * i];
        *im = pc->pw[2 * i + 1];
        return;
    }
    cpu_power_of_w(k, n, re, im);
}

static void FN(simplify)(long k, long n, long *ko, long *no) {
    k %= n;
    long g = gcd_l(k, n);
    if (g == 0) {
        *ko = 0;
        *no = 0;
        return;
    }
    k /= g;
    n /= g;
    if (2 * k > n) {
        k -= n;
    } else if (2 * k < -n) {
        k += n;
    }
    *ko = k;
    *no = n;
}

static void FN(esum)(FN(cv) * out, FN(cv) * const *x, int Nf, int dir, int kf, int nl,
                     FN(pair_cache) * pc) {
    for (int l = 0; l < nl; ++l) {
        out->re[l] = x[0]->re[l];
        out->im[l] = x[0]->im[l];
    }
    if (kf == 0) {
        for (int jf = 1; jf < Nf; ++jf) {
            for (int l = 0; l < nl; ++l) {
                out->re[l] += x[jf]->re[l];
                out->im[l] += x[jf]->im[l];
            }
        }
        return;
    }
    /* singletons and pairs in jf order */
    long sk[64], sn[64];
    int sj[64], ns = 0;
    long pk1[64], pn1[64], pk2[64], pn2[64];
    int pj1[64], pj2[64], np = 0;
    for (int jf = 1; jf < Nf; ++jf) {
        long k, n;
        FN(simplify)((long)dir * kf * jf, Nf, &k, &n);
        int found = -1;
        for (int s = 0; s < ns; ++s) {
            if (sk[s] == -k && sn[s] == n) {
                found = s;
                break;
            }
        }
        if (found >= 0) {
            pk1[np] = sk[found];
            pn1[np] = sn[found];
            pj1[np] = sj[found];
            pk2[np] = k;
            pn2[np] = n;
            pj2[np] = jf;
            ++np;
            for (int s = found; s + 1 < ns; ++s) {
                sk[s] = sk[s + 1];
                sn[s] = sn[s + 1];
                sj[s] = sj[s + 1];
            }
            --ns;
        } else {
            sk[ns] = k;
            sn[ns] = n;
            sj[ns] = jf;
            ++ns;
        }
    }
    for (int s = 0; s < ns; ++s) {
        double wr, wi;
        FN(power_of_w_t)(pc, sk[s], sn[s] ? sn[s] : 1, &wr, &wi);
        T r = (T)wr, i = (T)wi;
        for (int l = 0; l < nl; ++l) {
            T a = x[sj[s]]->re[l], b = x[sj[s]]->im[l];
            out->re[l] += a * r - b * i;
            out->im[l] += a * i + b * r;
        }
    }
    for (int p = 0; p < np; ++p) {
        /* reuse the conjugate pair computed for -kf */
        int hit = -1;
        for (int c = 0; c < pc->n; ++c) {
            if (pc->key[c][0] == -pk1[p] && pc->key[c][1] == pj1[p] &&
                pc->key[c][2] == -pk2[p] && pc->key[c][3] == pj2[p]) {
                hit = c;
                break;
            }
        }
        if (hit >= 0) {
            for (int l = 0; l < nl; ++l) {
                out->re[l] += pc->v1[hit].re[l] - pc->v2[hit].re[l];
                out->im[l] += pc->v1[hit].im[l] - pc->v2[hit].im[l];
            }
            continue;
        }
        long k1 = pk1[p], n1 = pn1[p];
        int j1 = pj1[p], j2 = pj2[p];
        if (k1 < 0) {
            k1 = pk2[p];
            n1 = pn2[p];
            int t = j1;
            j1 = j2;
            j2 = t;
        }
        double wr, wi;
        FN(power_of_w_t)(pc, k1, n1 ? n1 : 1, &wr, &wi);
        T r = (T)wr, i = (T)wi;
        int c = pc->n++;
        pc->key[c][0] = pk1[p];
        pc->key[c][1] = pj1[p];
        pc->key[c][2] = pk2[p];
        pc->key[c][3] = pj2[p];
        for (int l = 0; l < nl; ++l) {
            T x1r = x[j1]->re[l], x1i = x[j1]->im[l], x2r = x[j2]->re[l], x2i = x[j2]->im[l];
            pc->v1[c].re[l] = r * (x1r + x2r);
            pc->v1[c].im[l] = r * (x1i + x2i);
            pc->v2[c].re[l] = i * (x2i - x1i);
            pc->v2[c].im[l] = i * (x1r - x2r);
            out->re[l] += pc->v1[c].re[l] + pc->v2[c].re[l];
            out->im[l] += pc->v1[c].im[l] + pc->v2[c].im[l];
        }
    }
}

/* In-register in-place mixed-radix DFT (mixed_radix_fft.cpp:164-197); output in scrambled
 * order; the external twiddle (if any) multiplies the outputs of the last inner stage. */
static void FN(inplace)(FN(cv) * x, int nl, const int *fac, int L, int dir, const T *tw,
                        FN(pair_cache) * pcp) {
    int N = 1;
    for (int i = 0; i < L; ++i) {
        N *= fac[i];
    }
    int J = N, K = 1;
    FN(cv) y[64];
    FN(pair_cache) *pc = pcp;
    for (int f = L - 1; f >= 0; --f) {
        int Nf = fac[f];
        J /= Nf;
        for (int j = 0; j < J; ++j) {
            for (int k = 0; k < K; ++k) {
                FN(cv) *xp[64];
                for (int jf = 0; jf < Nf; ++jf) {
                    xp[jf] = &x[j + J * jf + J * Nf * k];
                }
                pc->n = 0;
                for (int kf = 0; kf < Nf; ++kf) {
                    FN(esum)(&y[kf], xp, Nf, dir, kf, nl, pc);
                    double wr, wi;
                    FN(power_of_w_t)(pc, (long)dir * kf * j, (long)J * Nf, &wr, &wi);
                    T r = (T)wr, i = (T)wi;
                    if (tw && f == 0) {
                        int idx = cpu_scramble(fac, L, j + J * kf + J * Nf * k, 0);
                        T tr = tw[2 * idx] * r - tw[2 * idx + 1] * i;
                        T ti = tw[2 * idx] * i + tw[2 * idx + 1] * r;
                        r = tr;
                        i = ti;
                    }
                    for (int l = 0; l < nl; ++l) {
                        T a = y[kf].re[l], b = y[kf].im[l];
                        y[kf].re[l] = a * r - b * i;
                        y[kf].im[l] = a * i + b * r;
                    }
                }
                for (int kf = 0; kf < Nf; ++kf) {
                    x[j + J * kf + J * Nf * k] = y[kf];
                }
            }
        }
        K *= Nf;
    }
}

/* One (M block, k) column tile of a plan: in/out interleaved complex arrays. */
static void FN(run_tile)(FN(cpu_plan) const *p, const T *in, T *out, int m0, int nl, long k,
                         FN(cv) * X1, FN(cv) * x, FN(pair_cache) * pc) {
    int const N = p->N;
    if (p->sbfft) {
        int fac[32];
        int L = trial_division_i(N, fac);
        for (int n = 0; n < N; ++n) {
            for (int l = 0; l < nl; ++l) {
                size_t o = (size_t)(m0 + l) + n * p->is1 + k * p->is2;
                x[n].re[l] = in[2 * o];
                x[n].im[l] = in[2 * o + 1];
            }
        }
        FN(inplace)(x, nl, fac, L, p->dir, NULL, pc);
        for (int n = 0; n < N; ++n) {
            int s = cpu_unscramble(fac, L, n, 0);
            for (int l = 0; l < nl; ++l) {
                size_t o = (size_t)(m0 + l) + n * p->os1 + k * p->os2;
                out[2 * o] = x[s].re[l];
                out[2 * o + 1] = x[s].im[l];
            }
        }
        return;
    }
    int J1 = N, J2 = 1, tw_offset = 0;
    for (int f = p->L - 1; f >= 0; --f) {
        int Nf = p->fac[f];
        J1 /= Nf;
        int ifac[32];
        int iL = trial_division_i(Nf, ifac);
        if (iL == 0) {
            ifac[0] = 1;
            iL = 1;
        }
        for (int j1j2 = 0; j1j2 < J1 * J2; ++j1j2) {
            int j1 = j1j2 % J1, j2 = j1j2 / J1;
            for (int i = 0; i < Nf; ++i) {
                if (f == p->L - 1) {
                    for (int l = 0; l < nl; ++l) {
                        size_t o = (size_t)(m0 + l) + (size_t)(j1 + J1 * i) * p->is1 + k * p->is2;
                        x[i].re[l] = in[2 * o];
                        x[i].im[l] = in[2 * o + 1];
                    }
                } else {
                    x[i] = X1[j1 + J1 * i + J1 * Nf * j2];
                }
            }
            const T *tw = f > 0 ? p->twiddle + 2 * (tw_offset + j1 * Nf) : NULL;
            FN(inplace)(x, nl, ifac, iL, p->dir, tw, pc);
            for (int i = 0; i < Nf; ++i) {
                X1[j1 + J1 * i + J1 * Nf * j2] = x[cpu_unscramble(ifac, iL, i, 0)];
            }
        }
        J2 *= Nf;
        tw_offset += J1 * Nf;
    }
    int Nf0 = p->fac[0];
    int JJ = N / Nf0;
    for (int j2 = 0; j2 < JJ; ++j2) {
        int u = cpu_unscramble(p->fac + 1, p->L - 1, j2, 1);
        for (int j1 = 0; j1 < Nf0; ++j1) {
            for (int l = 0; l < nl; ++l) {
                size_t o = (size_t)(m0 + l) + (size_t)(j2 + JJ * j1) * p->os1 + k * p->os2;
                out[2 * o] = X1[j1 + Nf0 * u].re[l];
                out[2 * o + 1] = X1[j1 + Nf0 * u].im[l];
            }
        }
    }
}

/* Plan for the reference's CPU device_info {8192, {4,8,16,32,64}, 32768, cpu}. */
static int FN(make_plan)(FN(cpu_plan) * p, int M, int N, int K, int dir, size_t is1, size_t is2,
                         size_t os1, size_t os2) {
    size_t const sizeof_real = sizeof(T);
    size_t const sgs = 4, sgs_max = 64, wg = 8192, lmem = 32768, reg_space = 32 * 64;
    memset(p, 0, sizeof(*p));
    p->N = N;
    p->M = M;
    p->K = K;
    p->dir = dir;
    p->is1 = is1;
    p->is2 = is2;
    p->os1 = os1;
    p->os2 = os2;
    size_t required = 2 * sizeof_real * (size_t)N * sgs;
    if (required < reg_space / 2) {
        p->sbfft = 1;
        size_t Mb = pow2_ge((size_t)M);
        if (Mb > sgs_max) {
            Mb = sgs_max;
        }
        p->Mb = (int)Mb;
        p->Kb = 1;
        return N <= 64 ? 0 : 1;
    }
    unsigned max_in_regs = (unsigned)((reg_space / 2) / (2 * sizeof_real) / sgs);
    unsigned f[MAXF];
    unsigned L = 2;
    for (unsigned index = 2; index <= 4; ++index) {
        cpu_factor((unsigned)N, index, f);
        L = index;
        unsigned fmax = 0;
        for (unsigned i = 0; i < index; ++i) {
            fmax = f[i] > fmax ? f[i] : fmax;
        }
        if (fmax < max_in_regs || is_prime_u(fmax)) {
            break;
        }
    }
    p->L = (int)L;
    unsigned fmax = 0;
    for (unsigned i = 0; i < L; ++i) {
        p->fac[i] = (int)f[i];
        fmax = f[i] > fmax ? f[i] : fmax;
    }
    if (fmax > 64) {
        return 1;
    }
    size_t Nb = pow2_ge((size_t)N / fmax);
    size_t max_Mb = wg / Nb;
    size_t slm_Mb = lmem / (2 * (size_t)N * sizeof_real);
    if (slm_Mb < max_Mb) {
        max_Mb = slm_Mb;
    }
    if (max_Mb >= sgs) {
        max_Mb = max_Mb / sgs * sgs;
    }
    if (max_Mb > sgs_max) {
        max_Mb = sgs_max;
    }
    size_t Mb = pow2_ge((size_t)M);
    p->Mb = (int)(Mb < max_Mb ? Mb : max_Mb);
    if (p->Mb < 1) {
        p->Mb = 1;
    }
    p->Kb = 1;
    /* twiddle table (factor2_slm_fft.hpp:49-90), double precision cast to T */
    int tw_size = 0, n = p->fac[0];
    for (int s = 1; s < p->L; ++s) {
        n *= p->fac[s];
        tw_size += n;
    }
    tw_size = 0;
    {
        int J1 = N;
        for (int s = p->L - 1; s >= 1; --s) {
            J1 /= p->fac[s];
            tw_size += J1 * p->fac[s];
        }
    }
    p->twiddle = (T *)malloc(sizeof(T) * 2 * (tw_size ? tw_size : 1));
    T *t = p->twiddle;
    int J1 = N;
    for (int s = p->L - 1; s >= 1; --s) {
        int Nf = p->fac[s];
        J1 /= Nf;
        for (int i = 0; i < J1; ++i) {
            for (int j = 0; j < Nf; ++j) {
                double arg = dir * 6.28318530717958647693 / (J1 * Nf) * i * j;
                t[2 * (j + Nf * i)] = (T)cos(arg);
                t[2 * (j + Nf * i) + 1] = (T)sin(arg);
            }
        }
        t += 2 * J1 * Nf;
    }
    return 0;
}

typedef struct {
    FN(cpu_plan) const *p;
    const T *in;
    T *out;
    long k0, k1;
} FN(job);

static void *FN(worker)(void *arg) {
    FN(job) *j = (FN(job) *)arg;
    FN(cpu_plan) const *p = j->p;
    int lanes = p->Mb > MAXLANES ? MAXLANES : p->Mb;
    FN(cv) *X1 = (FN(cv) *)malloc(sizeof(FN(cv)) * (size_t)p->N);
    FN(cv) *x = (FN(cv) *)malloc(sizeof(FN(cv)) * (size_t)(p->N > 64 ? p->N : 64));
    FN(pair_cache) *pc = (FN(pair_cache) *)malloc(sizeof(FN(pair_cache)));
    double *pw = (double *)malloc(sizeof(double) * 2 * (size_t)p->N);
    for (long i = 0; i < p->N; ++i) {
        cpu_power_of_w(i, p->N, &pw[2 * i], &pw[2 * i + 1]);
    }
    pc->pw = pw;
    pc->pwN = p->N;
    for (long k = j->k0; k < j->k1; ++k) {
        for (int m0 = 0; m0 < p->M; m0 += lanes) {
            int nl = p->M - m0 < lanes ? p->M - m0 : lanes;
            FN(run_tile)(p, j->in, j->out, m0, nl, k, X1, x, pc);
        }
    }
    free(pc);
    free(pw);
    free(X1);
    free(x);
    return NULL;
}

/* 1-D c2c of the M x N x K tensor (interleaved complex T) with the reference algorithm on
 * `threads` host threads. Returns 0 on success, 1 for unsupported sizes (radix > 64). */
int CAT(cpu_fft_c2c, SFX)(int M, int N, int K, int dir, size_t is1, size_t is2, size_t os1,
                          size_t os2, const T *in, T *out, int threads) {
    FN(cpu_plan) p;
    if (FN(make_plan)(&p, M, N, K, dir, is1, is2, os1, os2)) {
        free(p.twiddle);
        return 1;
    }
    if (threads < 1) {
        threads = 1;
    }
    if (threads > K) {
        threads = K;
    }
    pthread_t tid[256];
    FN(job) jobs[256];
    if (threads > 256) {
        threads = 256;
    }
    for (int t = 0; t < threads; ++t) {
        jobs[t].p = &p;
        jobs[t].in = in;
        jobs[t].out = out;
        jobs[t].k0 = (long)K * t / threads;
        jobs[t].k1 = (long)K * (t + 1) / threads;
        pthread_create(&tid[t], NULL, FN(worker), &jobs[t]);
    }
    for (int t = 0; t < threads; ++t) {
        pthread_join(tid[t], NULL);
    }
    free(p.twiddle);
    return 0;
}

/* Tiling the reference computes for its CPU device_info: {sbfft, L, fac[0..L-1], Mb}. */
int CAT(cpu_plan_info, SFX)(int M, int N, int *out) {
    FN(cpu_plan) p;
    int rc = FN(make_plan)(&p, M, N, 1, -1, 1, 1, 1, 1);
    out[0] = p.sbfft;
    out[1] = p.L;
    for (int i = 0; i < 4; ++i) {
        out[2 + i] = i < p.L ? p.fac[i] : 0;
    }
    out[6] = p.Mb;
    free(p.twiddle);
    return rc;
}

#undef FN
#undef CAT
#undef CAT_
