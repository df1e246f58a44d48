// The C++ drop-in API executed on the GPU, as a reference user writes it: configuration /
// parse_fft_descriptor, make_plan with a HIP stream and a jit_cache, plan.execute out of place,
// in place and with caller-owned events (the reference's Level Zero form), tensor_indexer for
// addressing. Known answers of the reference's tests (test/c2c.cpp:17-53 plane wave -> delta,
// test/r2c.cpp:40-56 cosine -> half deltas). Prints one line per case; exit status 0 = pass.
#include "bbfft/configuration.hpp"
#include "bbfft/hip/make_plan.hpp"
#include "bbfft/jit_cache_all.hpp"
#include "bbfft/parser.hpp"
#include "bbfft/tensor_indexer.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#define CHECK(x)                                                                                \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) {                                                                 \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);         \
            return 1;                                                                           \
        }                                                                                       \
    } while (0)

using namespace bbfft;
using cplx = std::complex<float>;

static double const tau = 6.283185307179586;

// c2c forward of x[m, n, k] = s e^{2 pi i b n / N} / N, b = (m + k) mod N: X = s delta(n - b)
static int c2c_case(char const *desc, bool inplace, bool events, hipStream_t stream,
                    jit_cache *cache) {
    configuration cfg = parse_fft_descriptor(desc);
    std::size_t const M = cfg.shape[0], N = cfg.shape[1], K = cfg.shape[2];
    // addressing of the (m, n, k) tensor through the public indexer (col-major: m fastest)
    using indexer = tensor_indexer<std::size_t, 3, layout::col_major>;
    indexer const xi({M, N, K}, {cfg.istride[0], cfg.istride[1], cfg.istride[2]});
    indexer const yi({M, N, K}, {cfg.ostride[0], cfg.ostride[1], cfg.ostride[2]});
    std::size_t const n_el = std::max(xi(M - 1, N - 1, K - 1), yi(M - 1, N - 1, K - 1)) + 1;
    std::vector<cplx> x(n_el), y(n_el);
    for (std::size_t k = 0; k < K; ++k) {
        for (std::size_t n = 0; n < N; ++n) {
            for (std::size_t m = 0; m < M; ++m) {
                std::size_t const b = (m + k) % N;
                double const ph = tau * static_cast<double>((b * n) % N) / static_cast<double>(N);
                double const s = (1.0 + static_cast<double>(k) / K) / static_cast<double>(N);
                x[xi(m, n, k)] =
                    cplx(static_cast<float>(s * std::cos(ph)), static_cast<float>(s * std::sin(ph)));
            }
        }
    }
    cplx *dx = nullptr, *dy = nullptr;
    CHECK(hipMalloc(&dx, n_el * sizeof(cplx)));
    CHECK(hipMemcpy(dx, x.data(), n_el * sizeof(cplx), hipMemcpyHostToDevice));
    if (!inplace) {
        CHECK(hipMalloc(&dy, n_el * sizeof(cplx)));
    }
    auto plan = make_plan(cfg, stream, cache);
    if (events) {
        hipEvent_t ready, done;
        CHECK(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
        CHECK(hipEventCreateWithFlags(&done, hipEventDisableTiming));
        CHECK(hipEventRecord(ready, stream));
        plan.execute(dx, inplace ? dx : dy, done, 1, &ready);
        CHECK(hipEventSynchronize(done));
        CHECK(hipEventDestroy(ready));
        CHECK(hipEventDestroy(done));
    } else {
        if (inplace) {
            plan.execute(dx);
        } else {
            plan.execute(dx, dy);
        }
        CHECK(hipStreamSynchronize(stream));
    }
    CHECK(hipMemcpy(y.data(), inplace ? dx : dy, n_el * sizeof(cplx), hipMemcpyDeviceToHost));
    double err = 0.0;
    for (std::size_t k = 0; k < K; ++k) {
        for (std::size_t n = 0; n < N; ++n) {
            for (std::size_t m = 0; m < M; ++m) {
                std::size_t const b = (m + k) % N;
                double const ref = n == b ? 1.0 + static_cast<double>(k) / K : 0.0;
                cplx const v = y[yi(m, n, k)];
                err = std::max(err, std::abs(std::complex<double>(v.real() - ref, v.imag())));
            }
        }
    }
    CHECK(hipFree(dx));
    if (dy) {
        CHECK(hipFree(dy));
    }
    bool const ok = err < 1e-4;
    std::printf("%s %s%s: max error %.3g %s\n", desc, inplace ? "in-place" : "out-of-place",
                events ? " (events)" : "", err, ok ? "ok" : "FAIL");
    return ok ? 0 : 1;
}

// r2c forward of x[n, k] = s cos(2 pi b n / N) / N, b = k mod N: X = s (delta(n-b) + delta(n+b))/2
static int r2c_case(char const *desc, hipStream_t stream, jit_cache *cache) {
    configuration cfg = parse_fft_descriptor(desc);
    cfg.set_strides_default(false);
    std::size_t const N = cfg.shape[1], K = cfg.shape[2], H = N / 2 + 1;
    std::vector<float> x(N * K);
    std::vector<cplx> y(H * K);
    for (std::size_t k = 0; k < K; ++k) {
        std::size_t const b = k % N;
        for (std::size_t n = 0; n < N; ++n) {
            double const s = (1.0 + static_cast<double>(k) / K) / static_cast<double>(N);
            x[n + k * N] = static_cast<float>(
                s * std::cos(tau * static_cast<double>((b * n) % N) / static_cast<double>(N)));
        }
    }
    float *dx = nullptr;
    cplx *dy = nullptr;
    CHECK(hipMalloc(&dx, x.size() * sizeof(float)));
    CHECK(hipMalloc(&dy, y.size() * sizeof(cplx)));
    CHECK(hipMemcpy(dx, x.data(), x.size() * sizeof(float), hipMemcpyHostToDevice));
    auto plan = make_plan(cfg, stream, cache);
    plan.execute(dx, dy);
    CHECK(hipStreamSynchronize(stream));
    CHECK(hipMemcpy(y.data(), dy, y.size() * sizeof(cplx), hipMemcpyDeviceToHost));
    double err = 0.0;
    for (std::size_t k = 0; k < K; ++k) {
        std::size_t const b = k % N;
        for (std::size_t n = 0; n < H; ++n) {
            double ref = 0.0;
            ref += n == b ? 0.5 : 0.0;
            ref += (n + b) % N == 0 ? 0.5 : 0.0;
            ref *= 1.0 + static_cast<double>(k) / K;
            cplx const v = y[n + k * H];
            err = std::max(err, std::abs(std::complex<double>(v.real() - ref, v.imag())));
        }
    }
    CHECK(hipFree(dx));
    CHECK(hipFree(dy));
    bool const ok = err < 1e-4;
    std::printf("%s: max error %.3g %s\n", desc, err, ok ? "ok" : "FAIL");
    return ok ? 0 : 1;
}

int main() {
    hipStream_t stream;
    if (hipStreamCreate(&stream) != hipSuccess) {
        std::printf("no HIP device\n");
        return 2;
    }
    jit_cache_all cache;
    int failures = 0;
    try {
        failures += c2c_case("scfo8.256*512", false, false, stream, &cache);
        failures += c2c_case("scfi8.256*512", true, false, stream, &cache);
        failures += c2c_case("scfo3.60*77", false, true, stream, &cache);
        failures += c2c_case("scfo8.256*512", false, true, stream, &cache); // cached kernel
        // padded, non-packed strides in and out (M = 4 padded to 5, N to 18 rows)
        failures += c2c_case("scfo4.16*8i1,5,90o1,6,108", false, false, stream, &cache);
        failures += r2c_case("srfo128*300", stream, &cache);
        failures += r2c_case("srfo27*33", stream, &cache);
    } catch (std::exception const &e) {
        std::printf("exception: %s\n", e.what());
        return 1;
    }
    if (hipStreamDestroy(stream) != hipSuccess) {
        return 1;
    }
    std::printf("%s\n", failures ? "FAILED" : "PASSED");
    return failures ? 1 : 0;
}
