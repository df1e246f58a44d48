"""Pins of the oracle (CPU): the FP64 DFT against a library routine (numpy's pocketfft), the
reference's analytic known answers, Parseval and linearity; the CPU port of the reference
algorithm against the FP64 DFT, the reference's factor/scrambler tables and the reference's
own integer sources (oracle/_ref)."""
import ctypes
import os

import numpy as np
import pytest

import oracle
import oracle.cpu_port as port

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref.so")
TAU = 2 * np.pi


def _packed(shape):
    s, out = 1, []
    for n in shape:
        out.append(s)
        s *= n
    return out + [0] * (5 - len(out))


@pytest.mark.parametrize("N", [1, 2, 3, 5, 8, 12, 17, 64, 105, 256])
@pytest.mark.parametrize("d", [-1, 1])
def test_c2c_matches_pocketfft(N, d):
    M, K = 3, 4
    x = oracle.uniform_complex(N, M * N * K)
    st = _packed([M, N, K])
    y = oracle.fft(1, [M, N, K], st, st, 0, d, x)
    xa = x.reshape(K, N, M)
    ref = np.fft.fft(xa, axis=1) if d < 0 else np.fft.ifft(xa, axis=1) * N
    assert oracle.rel_l2(y, ref.ravel()) < 1e-14


def test_c2c_strided_and_nd_match_pocketfft():
    # non-packed strides {1, M+1, (M+1)(N+1)} (reference test/c2c.cpp:75)
    M, N, K = 2, 6, 3
    st = [1, M + 1, (M + 1) * (N + 1), 0, 0]
    x = np.zeros(st[2] * K, complex)
    xs = oracle.uniform_complex(3, M * N * K).reshape(K, N, M)
    for k in range(K):
        for n in range(N):
            for m in range(M):
                x[m + n * st[1] + k * st[2]] = xs[k, n, m]
    y = oracle.fft(1, [M, N, K], st, st, 0, -1, x)
    ref = np.fft.fft(xs, axis=1)
    got = np.array([[[y[m + n * st[1] + k * st[2]] for m in range(M)] for n in range(N)]
                    for k in range(K)])
    assert oracle.rel_l2(got, ref) < 1e-14
    shape = [2, 5, 6, 7, 3]
    x = oracle.uniform_complex(4, int(np.prod(shape)))
    y = oracle.fft(3, shape, _packed(shape), _packed(shape), 0, -1, x)
    ref = np.fft.fftn(x.reshape(shape[::-1]), axes=(1, 2, 3))
    assert oracle.rel_l2(y, ref.ravel()) < 1e-14


@pytest.mark.parametrize("dims", [[5], [8], [10, 3], [4, 8, 2], [5, 4, 6]])
def test_r2c_and_c2r_match_pocketfft(dims):
    M, K = 2, 3
    shape = [M] + dims + [K]
    half = [M, dims[0] // 2 + 1] + dims[1:] + [K]
    D = len(dims)
    x = oracle.uniform_real(6, int(np.prod(shape)))
    axes = tuple(range(1, D + 1))
    xa = x.reshape(shape[::-1])
    X = oracle.fft(D, shape, _packed(shape), _packed(half), 1, -1, x)
    # numpy array axes are the modes in reverse order: axis D is mode 1, which rfftn halves as
    # the last axis listed
    Xr = np.fft.rfftn(xa, axes=axes)
    assert oracle.rel_l2(X, Xr.ravel()) < 1e-14
    # c2r of a Hermitian-consistent spectrum (DC and Nyquist imaginary parts zero)
    xb = oracle.fft(D, shape, _packed(half), _packed(shape), 2, 1, Xr.ravel())
    assert oracle.rel_l2(xb, x * np.prod(dims)) < 1e-14


def test_plane_wave_known_answer():
    """reference test/c2c.cpp:28-47: x = s(k) e^{2 pi i b n / N} / N -> X = s(k) delta(n - b)."""
    M, N, K = 3, 12, 4
    m, n, k = np.meshgrid(range(M), range(N), range(K), indexing="ij")
    b = (m + k) % N
    s = 1 + k / K
    x = np.zeros(M * N * K, complex)
    off = (m + M * n + M * N * k).ravel()
    x[off] = (s * np.exp(1j * TAU * b * n / N) / N).ravel()
    y = oracle.fft(1, [M, N, K], _packed([M, N, K]), _packed([M, N, K]), 0, -1, x)
    # inputs are rounded to double, so agreement is to ~N eps
    assert np.allclose(y[off], (s * ((n - b) % N == 0)).ravel(), rtol=0, atol=1e-13)


def test_cosine_known_answer_r2c_and_c2r():
    """reference test/r2c.cpp:40-131: cosine <-> half-deltas at +-b."""
    M, N, K = 2, 10, 3
    m, n, k = np.meshgrid(range(M), range(N), range(K), indexing="ij")
    b = (m + k) % N
    s = 1 + k / K
    x = np.zeros(M * N * K)
    off = (m + M * n + M * N * k).ravel()
    x[off] = (s * np.cos(TAU * b * n / N) / N).ravel()
    H = N // 2 + 1
    X = oracle.fft(1, [M, N, K], _packed([M, N, K]), _packed([M, H, K]), 1, -1, x)
    mh, nh, kh = np.meshgrid(range(M), range(H), range(K), indexing="ij")
    bh = (mh + kh) % N
    ref = (1 + kh / K) * ((((nh - bh) % N) == 0).astype(float)
                         + (((nh + bh) % N) == 0).astype(float)) / 2
    offh = (mh + M * nh + M * H * kh).ravel()
    assert np.allclose(X[offh], ref.ravel(), rtol=0, atol=1e-13)
    xb = oracle.fft(1, [M, N, K], _packed([M, H, K]), _packed([M, N, K]), 2, 1, X)
    assert np.allclose(xb[off], (s * np.cos(TAU * b * n / N)).ravel(), rtol=0, atol=1e-13)


def test_c2r_ignores_dc_imag_and_applies_nyquist_rule():
    """imag(X[0]) is ignored (reference test/r2c.cpp:310-324); for even N the reference's
    half-length pre-pass adds -imag(X[N/2]) to every sample (src/base/generator/sbfft_gen.cpp:
    225-246; pinned independently by the restatement tests below)."""
    for N in (7, 8):
        H = N // 2 + 1
        X = oracle.uniform_complex(8, H)
        base = oracle.fft(1, [1, N, 1], [1, 1, N], [1, 1, N], 2, 1, X)
        Xp = X.copy()
        Xp[0] += 5j
        assert np.array_equal(oracle.fft(1, [1, N, 1], [1, 1, N], [1, 1, N], 2, 1, Xp), base)
        if N % 2 == 0:
            Xn = X.copy()
            Xn[N // 2] += 0.75j
            diff = oracle.fft(1, [1, N, 1], [1, 1, N], [1, 1, N], 2, 1, Xn) - base
            assert np.allclose(diff, -0.75, atol=1e-14)


def _ref_c2r_even(X, N):
    """numpy restatement of the reference's even-length c2r, independent of the oracle:
    the half-length pre-pass of src/base/generator/sbfft_gen.cpp:226-247 (f2fft_gen.cpp:350-366
    is the same per i) for i = 0..N/4 -- imag(X[0]) zeroed, x1 = X[i], x2 = conj(X[N/2 - i]),
    a = x1 + x2, b = (x1 - x2) * i w_N^i, z[i mod N/2] = a + b and, when the slots differ,
    z[(N/2 - i) mod N/2] = (a.re - b.re, b.im - a.im) -- then the unnormalised N/2-point backward
    FFT (numpy's pocketfft) and the de-interleave x[2n] = re z[n], x[2n+1] = im z[n]
    (sbfft_gen.cpp:201-212, reshaped_mode(0, {2, N_fft}))."""
    H = N // 2
    z = np.zeros(H, complex)
    for i in range(N // 4 + 1):
        io = H - i
        x1 = X[i]
        if i == 0:
            x1 = complex(x1.real, 0.0)
        x2 = np.conj(X[io])
        a = x1 + x2
        b = (x1 - x2) * (np.exp(2j * np.pi * i / N) * 1j)
        z[i % H] = a + b
        if i % H != io % H:
            z[io % H] = complex(a.real - b.real, b.imag - a.imag)
    y = np.fft.ifft(z) * H
    x = np.empty(N)
    x[0::2] = y.real
    x[1::2] = y.imag
    return x


def _ref_c2r_odd_pair(A, B, N):
    """numpy restatement of the reference's odd-length c2r, which packs two K columns into one
    complex FFT (src/base/generator/sbfft_gen.cpp:318-351): z[0] = (re A[0], re B[0]),
    z[i] = A[i] + i B[i] and z[N - i] = conj(A[i]) + i conj(B[i]) for i = 1..N/2, then the
    unnormalised N-point backward FFT; column a is the real part, column b the imaginary part."""
    z = np.zeros(N, complex)
    z[0] = complex(A[0].real, B[0].real)
    for i in range(1, N // 2 + 1):
        z[i] = A[i] + 1j * B[i]
        z[N - i] = np.conj(A[i]) + 1j * np.conj(B[i])
    y = np.fft.ifft(z) * N
    return y.real, y.imag


@pytest.mark.parametrize("N", [2, 4, 6, 8, 16, 48, 96, 128, 256, 300, 512])
def test_c2r_even_matches_reference_prepass_on_non_hermitian_spectra(N):
    """Fully random half spectra (non-zero imag(X[0]) and imag(X[N/2])): the oracle's c2r agrees
    with the restated reference pre-pass. N/2 is odd for N = 2 and 6 (the i = N/4 slot is not
    self-paired) and even for the other lengths."""
    M, K = 3, 4
    H = N // 2 + 1
    X = oracle.uniform_complex(100 + N, M * H * K)
    got = oracle.fft(1, [M, N, K], _packed([M, H, K]), _packed([M, N, K]), 2, 1, X)
    Xa = X.reshape(K, H, M)
    ref = np.empty((K, N, M))
    for k in range(K):
        for m in range(M):
            ref[k, :, m] = _ref_c2r_even(Xa[k, :, m], N)
    assert oracle.rel_l2(got, ref.ravel()) <= 1e-13


@pytest.mark.parametrize("N", [1, 3, 5, 7, 15, 63, 105, 243])
def test_c2r_odd_matches_reference_pair_packing(N):
    M, K = 2, 6
    H = N // 2 + 1
    X = oracle.uniform_complex(200 + N, M * H * K)
    got = oracle.fft(1, [M, N, K], _packed([M, H, K]), _packed([M, N, K]), 2, 1, X)
    Xa = X.reshape(K, H, M)
    ref = np.empty((K, N, M))
    for k in range(0, K, 2):
        for m in range(M):
            ref[k, :, m], ref[k + 1, :, m] = _ref_c2r_odd_pair(Xa[k, :, m], Xa[k + 1, :, m], N)
    assert oracle.rel_l2(got, ref.ravel()) <= 1e-13


@pytest.mark.parametrize("dims", [[8, 6], [6, 5, 4], [16, 3], [7, 4, 3], [12, 10, 2]])
def test_c2r_nd_matches_reference_order_on_non_hermitian_spectra(dims):
    """Multi-dimensional c2r in the reference's order (src/common/algorithm/nd_fft.hpp:89-94):
    complex backward transforms over modes 2..d (numpy), then the real transform over mode 1 as
    restated above, on fully random half spectra."""
    M, K = 2, 2
    N1 = dims[0]
    H = N1 // 2 + 1
    half = [M, H] + dims[1:] + [K]
    shape = [M] + dims + [K]
    D = len(dims)
    X = oracle.uniform_complex(300 + N1, int(np.prod(half)))
    got = oracle.fft(D, shape, _packed(half), _packed(shape), 2, 1, X)
    # numpy axes are the modes reversed: axis 0 = K, axis D + 1 = M, mode d = axis D + 1 - d
    Xa = X.reshape(half[::-1])
    for d in range(2, D + 1):
        ax = D + 1 - d
        Xa = np.fft.ifft(Xa, axis=ax) * Xa.shape[ax]
    Xa = np.moveaxis(Xa, D, -1)  # mode 1 last
    out = np.empty(Xa.shape[:-1] + (N1,))
    for idx in np.ndindex(*Xa.shape[:-1]):
        if N1 % 2 == 0:
            out[idx] = _ref_c2r_even(Xa[idx], N1)
    if N1 % 2 == 1:
        # pair columns along K (axis 0), as the reference packs consecutive K columns
        for idx in np.ndindex(*Xa.shape[1:-1]):
            for k in range(0, K, 2):
                out[(k,) + idx], out[(k + 1,) + idx] = _ref_c2r_odd_pair(
                    Xa[(k,) + idx], Xa[(k + 1,) + idx], N1)
    ref = np.moveaxis(out, -1, D)
    assert oracle.rel_l2(got, ref.ravel()) <= 1e-13


def test_c2r_restatement_detects_planted_nyquist_mistakes():
    """The restatement separates the oracle's rule from the two plausible mistakes: dropping
    imag(X[N/2]) or flipping its sign (both planted on the CPU, in a copy of the input)."""
    N = 16
    H = N // 2 + 1
    X = oracle.uniform_complex(41, H)
    ref = _ref_c2r_even(X, N)
    dropped = X.copy()
    dropped[N // 2] = dropped[N // 2].real
    flipped = X.copy()
    flipped[N // 2] = np.conj(flipped[N // 2])
    assert oracle.rel_l2(_ref_c2r_even(dropped, N), ref) > 1e-3
    assert oracle.rel_l2(_ref_c2r_even(flipped, N), ref) > 1e-3
    got = oracle.fft(1, [1, N, 1], [1, 1, N], [1, 1, N], 2, 1, X)
    assert oracle.rel_l2(got, ref) <= 1e-13


def test_parseval_and_linearity():
    M, N, K = 4, 60, 5
    st = _packed([M, N, K])
    x = oracle.uniform_complex(12, M * N * K)
    z = oracle.uniform_complex(13, M * N * K)
    X = oracle.fft(1, [M, N, K], st, st, 0, -1, x)
    Z = oracle.fft(1, [M, N, K], st, st, 0, -1, z)
    assert abs(np.vdot(X, X).real / (N * np.vdot(x, x).real) - 1) < 1e-14
    W = oracle.fft(1, [M, N, K], st, st, 0, -1, 2.5 * x - 1j * z)
    assert oracle.rel_l2(W, 2.5 * X - 1j * Z) < 1e-14


def test_seeded_inputs_are_reproducible_and_uniform():
    a = oracle.uniform(7, 100000)
    assert np.array_equal(a, oracle.uniform(7, 100000))
    assert np.array_equal(oracle.uniform(7, 10, offset=50), a[50:60])
    assert -1 <= a.min() and a.max() < 1 and abs(a.mean()) < 0.01
    assert not np.array_equal(a, oracle.uniform(8, 100000))


# ---- the CPU port of the reference algorithm -------------------------------------------

@pytest.mark.parametrize("precision", [4, 8])
@pytest.mark.parametrize("N", [2, 3, 5, 13, 16, 17, 27, 32, 60, 63, 64, 105, 128, 243, 256, 363,
                               384, 512])
def test_port_matches_fp64_dft(N, precision):
    M, K = 3, 5
    x = oracle.uniform_complex(N + precision, M * N * K)
    for d in (-1, 1):
        y = port.fft_c2c(x, M, N, K, d, precision, threads=2)
        st = _packed([M, N, K])
        ref = oracle.fft(1, [M, N, K], st, st, 0, d, x)
        assert oracle.rel_l2(y, ref) <= oracle.tolerance(precision, N)


@pytest.mark.parametrize("n,index,expect", [
    (4096, 3, [16, 16, 16]), (2080, 4, [4, 5, 8, 13]), (3465, 3, [11, 15, 21]),
    (9216, 4, [8, 8, 12, 12]), (64536, 3, [1, 24, 2689]), (65536, 3, [32, 32, 64])])
def test_port_factor_reference_table(n, index, expect):
    assert port.factor(n, index) == expect  # reference test/codegen.cpp:71-125


def test_port_scrambler_reference_table():
    """reference test/codegen.cpp:15-69."""
    fac = [3, 2, 2]
    table = [0, 4, 8, 2, 6, 10, 1, 5, 9, 3, 7, 11]
    for i, e in enumerate(table):
        assert port.scramble(fac, i, True) == e
        assert port.unscramble(fac, e, True) == i
    assert port.scramble(fac, 12) == 12 and port.scramble(fac, 13) == 16
    assert port.unscramble(fac, 12) == 12 and port.unscramble(fac, 16) == 13
    fac = [13, 5, 2, 7]
    for i in range(2 * 910):
        assert port.scramble(fac, port.unscramble(fac, i)) == i
        assert port.unscramble(fac, port.scramble(fac, i)) == i


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="reference sources not built")
def test_port_integer_logic_matches_reference_build():
    ref = ctypes.CDLL(REF_LIB)
    out = (ctypes.c_uint * 4)()
    for index in (2, 3, 4):
        for n in range(1, 1100):
            ref.ref_factor(n, index, out)
            assert port.factor(n, index) == list(out)[:index]
    rng = np.random.default_rng(0)
    for _ in range(200):
        fac = [int(v) for v in rng.integers(2, 9, size=int(rng.integers(1, 5)))]
        arr = (ctypes.c_int * len(fac))(*fac)
        n = int(np.prod(fac))
        for i in range(0, 2 * n, max(1, n // 7)):
            for in0 in (0, 1):
                if in0 and i >= n:
                    continue
                assert port.scramble(fac, i, in0) == ref.ref_scramble(arr, len(fac), i, in0)
                assert port.unscramble(fac, i, in0) == ref.ref_unscramble(arr, len(fac), i, in0)


def test_port_power_of_w_reference_table():
    """power_of_w for n <= 64 reproduces the reference's correctly rounded tables."""
    import json
    data = json.load(open(os.path.join(ROOT, "tests", "golden", "w_lut.json")))["values"]
    for n, vals in data.items():
        for i, (re_, im_) in enumerate(vals):
            w = port.power_of_w(i, int(n))
            assert (w.real, w.imag) == (float.fromhex(re_), float.fromhex(im_)), (n, i)


def test_port_reference_tilings():
    """Dispatch and factorisations of the reference's CPU device_info."""
    assert port.plan_info(8, 16, 4)["sbfft"] == 1
    assert port.plan_info(8, 256, 4)["factorization"] == [16, 16]
    assert port.plan_info(8, 512, 4)["factorization"] == [8, 8, 8]
    assert port.plan_info(8, 60, 4)["factorization"] == [6, 10]
    assert port.plan_info(8, 384, 4)["factorization"] == [16, 24]
    assert port.plan_info(8, 16, 8)["sbfft"] == 0
