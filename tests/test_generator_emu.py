"""The generated HIP kernels, compiled for the host and run by the fiber emulator
(tests/emu/hip_emu.hpp), against the FP64 oracle. CPU only: covers the generator's indexing,
LDS exchange, barriers, tile loop, real pre/post passes, callbacks and nd pass chaining."""
import numpy as np
import pytest

import bbfft_amd as bb
import callback_sources as cbs
import oracle
from emu import runner


def _cfg(M, N, K, fp=4, d=-1, t=0, inplace=False, dims=None):
    dims = dims or [N]
    cfg = bb.configuration(len(dims), [M] + dims + [K], bb.precision(fp), bb.direction(d),
                           bb.transform_type(t))
    cfg.set_strides_default(inplace)
    return cfg


@pytest.mark.parametrize("fp", [4, 8])
@pytest.mark.parametrize("N", [2, 3, 7, 8, 12, 16, 17, 32, 60, 64, 105, 128, 243, 256, 384, 512])
def test_c2c(N, fp):
    for M, K, d, cap in [(3, 5, -1, 2), (8, 3, 1, None)]:
        cfg = _cfg(M, N, K, fp, d)
        x = oracle.uniform_complex(N, M * N * K)
        y = runner.run(bb, cfg, x, M * N * K, grid_cap=cap)
        ref = oracle.fft(1, cfg.shape, cfg.istride, cfg.ostride, 0, d, x)
        assert oracle.rel_l2(y, ref) <= oracle.tolerance(fp, N)


@pytest.mark.parametrize("N", [2, 5, 8, 10, 26, 27, 64, 102, 105, 220, 256])
@pytest.mark.parametrize("inplace", [False, True])
def test_r2c(N, inplace):
    for M, K, fp in [(1, 5, 4), (3, 4, 8)]:
        cfg = _cfg(M, N, K, fp, -1, 1, inplace)
        n_in = oracle.extent(1, cfg.shape, cfg.istride)
        n_out = oracle.extent(1, [M, N // 2 + 1, K], cfg.ostride)
        x = oracle.uniform_real(N, n_in)
        ref = oracle.fft(1, cfg.shape, cfg.istride, cfg.ostride, 1, -1, x, out_size=n_out)
        if inplace:
            rt = np.float32 if fp == 4 else np.float64
            buf = np.zeros(max(n_in, 2 * n_out), rt)
            buf[:n_in] = x
            y = runner.run(bb, cfg, None, 0, in_place_buffer=buf).view(
                np.complex64 if fp == 4 else np.complex128)[:n_out]
        else:
            y = runner.run(bb, cfg, x, n_out, grid_cap=2)
        assert oracle.rel_l2(y, ref) <= oracle.tolerance(fp, N)


@pytest.mark.parametrize("N", [2, 3, 11, 16, 21, 25, 48, 96, 300, 315, 512])
@pytest.mark.parametrize("inplace", [False, True])
def test_c2r(N, inplace):
    """DC and Nyquist imaginary parts are non-zero in the inputs (reference rules)."""
    for M, K, fp in [(1, 5, 4), (5, 3, 8)]:
        cfg = _cfg(M, N, K, fp, 1, 2, inplace)
        n_in = oracle.extent(1, [M, N // 2 + 1, K], cfg.istride)
        n_out = oracle.extent(1, cfg.shape, cfg.ostride)
        X = oracle.uniform_complex(N, n_in)
        ref = oracle.fft(1, cfg.shape, cfg.istride, cfg.ostride, 2, 1, X, out_size=n_out)
        if inplace:
            ct = np.complex64 if fp == 4 else np.complex128
            buf = np.zeros(max(n_in, (n_out + 1) // 2), ct)
            buf[:n_in] = X
            y = runner.run(bb, cfg, None, 0, in_place_buffer=buf).view(
                np.float32 if fp == 4 else np.float64)[:n_out]
        else:
            y = runner.run(bb, cfg, X, n_out, grid_cap=2)
        m, n, k = np.meshgrid(range(M), range(N), range(K), indexing="ij")
        sel = (m + n * cfg.ostride[1] + k * cfg.ostride[2]).ravel()
        assert oracle.rel_l2(y[sel], ref[sel]) <= oracle.tolerance(fp, N)


@pytest.mark.parametrize("dims,t", [([8, 5], 0), ([5, 6, 4], 0), ([8, 6], 1), ([5, 4, 6], 1),
                                    ([8, 6], 2), ([5, 4, 6], 2)])
def test_nd_pass_chain(dims, t):
    """2-D/3-D transforms as the plan chains them: pass 0 in->out, later passes out->out."""
    M, K = 2, 3
    cfg = _cfg(M, 0, K, 4, -1 if t != 2 else 1, t, False, dims)
    shape = cfg.shape
    half = list(shape)
    half[1] = dims[0] // 2 + 1
    D = len(dims)
    ishape, oshape = (shape, half) if t == 1 else (half, shape) if t == 2 else (shape, shape)
    n_in = oracle.extent(D, ishape, cfg.istride)
    n_out = oracle.extent(D, oshape, cfg.ostride)
    x = oracle.uniform_real(3, n_in) if t == 1 else oracle.uniform_complex(3, n_in)
    ref = oracle.fft(D, shape, cfg.istride, cfg.ostride, t, int(cfg.dir), x, out_size=n_out)
    assert bb.num_passes(cfg) == D
    cur = runner.run(bb, cfg, x, n_out, pass_index=0)
    for p in range(1, D):
        cur = runner.run(bb, cfg, None, 0, pass_index=p, in_place_buffer=cur) if p < D - 1 \
            else runner.run(bb, cfg, cur, n_out, pass_index=p)
    assert oracle.rel_l2(cur, ref) <= oracle.tolerance(4, int(np.prod(dims)))


@pytest.mark.parametrize("fp,M,N", [(4, 1, 8), (4, 4, 64), (8, 3, 212)])
def test_callbacks_bitwise(fp, M, N):
    """Zero-padding load and truncating/scaling store callbacks equal the plain plans bitwise
    (reference test/callback.cpp)."""
    K = 4
    real_t = "float" if fp == 4 else "double"
    N_ext = 2 * N
    n_ref, n_have = N_ext // 2 + 1, N // 2 + 1
    plain = bb.configuration(1, [M, N_ext, K], bb.precision(fp), bb.direction.backward,
                             bb.transform_type.c2r, [1, M, M * n_ref], [1, M, M * N_ext])
    cb = bb.configuration(1, [M, N_ext, K], bb.precision(fp), bb.direction.backward,
                          bb.transform_type.c2r, [1, M, M * n_ref], [1, M, M * N_ext],
                          bb.user_module(cbs.load_hip(real_t, M, n_ref, n_have), "load"))
    X = oracle.uniform_complex(5, M * n_have * K).reshape(K, n_have, M)
    Xp = np.zeros((K, n_ref, M), complex)
    Xp[:, :n_have] = X
    compact = np.zeros(Xp.size, complex)
    compact[:X.size] = X.ravel()
    a = runner.run(bb, plain, Xp.ravel(), M * N_ext * K)
    b = runner.run(bb, cb, compact, M * N_ext * K)
    assert np.array_equal(a, b)
    # store: keep N/4 bins scaled by 1/N, HIP and OpenCL-C dialects
    n_h, n_cut = N // 2 + 1, N // 4
    scale = float(np.float32(1.0 / N)) if fp == 4 else 1.0 / N
    plain = bb.configuration(1, [M, N, K], bb.precision(fp), bb.direction.forward,
                             bb.transform_type.r2c, [1, M, M * N], [1, M, M * n_h])
    x = oracle.uniform_real(6, M * N * K)
    ref = runner.run(bb, plain, x, M * n_h * K).reshape(K, n_h, M)
    rt = np.float32 if fp == 4 else np.float64
    for src, lang in [(cbs.store_hip, bb.kernel_language.hip),
                      (cbs.store_opencl, bb.kernel_language.opencl_c)]:
        cfg = bb.configuration(1, [M, N, K], bb.precision(fp), bb.direction.forward,
                               bb.transform_type.r2c, [1, M, M * N], [1, M, M * n_h],
                               bb.user_module(src(real_t, M, n_h, n_cut, scale.hex()), None,
                                              "store", lang))
        y = runner.run(bb, cfg, x, M * n_h * K)[:M * n_cut * K].reshape(K, n_cut, M)
        assert np.array_equal(y.real, ref.real[:, :n_cut] * rt(scale))
        assert np.array_equal(y.imag, ref.imag[:, :n_cut] * rt(scale))


def _real_inner(a, b):
    return float(np.vdot(np.asarray(a, np.complex128).ravel(),
                         np.asarray(b, np.complex128).ravel()).real)


@pytest.mark.parametrize("N", [8, 9, 16, 15, 60, 105, 256])
@pytest.mark.parametrize("t", [0, 1, 2])
def test_adjoint_kernels_inner_product_identity(N, t):
    """<A x, y> = <x, A^T y> (real inner product) for the forward kernel A and the adjoint
    kernel (opposite-sign c2c; Hermitian-fold c2r for r2c; doubling r2c for c2r)."""
    M, K, fp = 3, 5, 8
    d = 1 if t == 2 else -1
    cfg = _cfg(M, N, K, fp, d, t, False)
    adj = bb.adjoint_configuration(cfg)
    H = N // 2 + 1
    ishape = [M, H if t == 2 else N, K]
    oshape = [M, H if t == 1 else N, K]
    n_in = oracle.extent(1, ishape, cfg.istride)
    n_out = oracle.extent(1, oshape, cfg.ostride)
    x = oracle.uniform_real(1, n_in) if t == 1 else oracle.uniform_complex(1, n_in)
    y = oracle.uniform_real(2, n_out) if t == 2 else oracle.uniform_complex(2, n_out)
    Ax = runner.run(bb, cfg, x, n_out)
    variant = {"flags": 4} if t != 0 else {"flags": 0}
    ATy = runner.run(bb, adj, y, n_in, tuning=variant)
    lhs, rhs = _real_inner(Ax, y), _real_inner(x, ATy)
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + np.linalg.norm(Ax) * np.linalg.norm(y))


@pytest.mark.parametrize("fp", [4, 8])
@pytest.mark.parametrize("t,N", [(0, 16), (0, 60), (0, 256), (0, 105), (1, 64), (1, 27), (1, 10)])
@pytest.mark.parametrize("cap", [1, 2, 3])
def test_prefetch_tile_loop(t, N, fp, cap):
    """Software-pipelined tile loop (fft1d flag 64): with `cap` work-groups every group walks
    several tiles, loading the next tile's inputs while it transforms the current one; edge
    tiles (M not a multiple of Mb, odd K for the paired odd-N real kernels) included."""
    for M, K, inplace in [(3, 7, False), (5, 4, True), (1, 9, False)]:
        cfg = _cfg(M, N, K, fp, -1, t, inplace)
        # in-place real transforms need the whole M in one tile (the reference's rule)
        tuning = {"flags": bb.PREFETCH, "Mb": M if (t and inplace) else min(M, 2), "Kb": 1}
        kc = bb.kernel_configuration(cfg, tuning=tuning)
        assert kc["name"].endswith("_g64")
        n_in = oracle.extent(1, cfg.shape, cfg.istride)
        n_out = oracle.extent(1, [M, N // 2 + 1 if t else N, K], cfg.ostride)
        x = oracle.uniform_real(N, n_in) if t else oracle.uniform_complex(N, n_in)
        ref = oracle.fft(1, cfg.shape, cfg.istride, cfg.ostride, t, -1, x, out_size=n_out)
        if inplace:
            rt = np.float32 if fp == 4 else np.float64
            ct = np.complex64 if fp == 4 else np.complex128
            buf = np.zeros(max(n_in, 2 * n_out) if t else 2 * max(n_in, n_out), rt)
            if t:
                buf[:n_in] = x
            else:
                buf.view(ct)[:n_in] = x
            y = runner.run(bb, cfg, None, 0, in_place_buffer=buf, tuning=tuning,
                           grid_cap=cap).view(ct)[:n_out]
        else:
            y = runner.run(bb, cfg, x, n_out, grid_cap=cap, tuning=tuning)
        assert oracle.rel_l2(y, ref) <= oracle.tolerance(fp, N)


def test_prefetch_rejected_where_unsupported():
    with pytest.raises(bb.BadConfiguration):
        bb.kernel_configuration(_cfg(2, 16, 3, 4, 1, 2), tuning={"flags": bb.PREFETCH})


_ND_SCALE_SRC = """
__device__ float2 ld(float2* in, size_t offset) {
    float2 v = in[offset]; v.x *= 0.125f; v.y *= 0.125f; return v;
}
__device__ float ldr(float* in, size_t offset) { return in[offset] * 0.125f; }
__device__ void st(float2* out, size_t offset, float2 v) {
    v.x *= 4.0f; v.y *= 4.0f; out[offset] = v;
}
__device__ void str_(float* out, size_t offset, float v) { out[offset] = v * 4.0f; }
"""


def _run_nd_passes(cfg, x):
    """The plan's 1-D passes in execution order, each through the emulator (out of place)."""
    cur = x
    for i in range(bb.num_passes(cfg)):
        kc = bb.kernel_configuration(cfg, i)
        t = runner.pass_type(cfg, i)
        n_out = kc["M"] * (kc["N"] // 2 + 1 if t == 1 else kc["N"]) * kc["K"]
        cur = runner.run(bb, cfg, cur, n_out, pass_index=i)
    return cur


@pytest.mark.parametrize("desc", ["scfo16x8x4*3", "srfo16x8x4*3", "srbo16x8x4*3",
                                  "scbo2.12x10*5"])
def test_nd_callbacks_first_and_last_pass(desc):
    """Callbacks on 2-D/3-D plans (an extension: the reference takes them on 1-D plans only):
    the load callback runs in the first pass and sees the plan's input offsets, the store
    callback in the last pass and sees the plan's output offsets. Scaling by powers of two is
    exact, so the result equals the plain plan's with the input scaled by 1/8 and the output
    by 4, bit for bit."""
    cfg = bb.parse_fft_descriptor(desc)
    t = int(cfg.type)
    plain = bb.parse_fft_descriptor(desc)
    cfg.callbacks = bb.user_module(_ND_SCALE_SRC, "ldr" if t == 1 else "ld",
                                   "str_" if t == 2 else "st", bb.kernel_language.hip)
    names = [bb.kernel_configuration(cfg, i)["name"] for i in range(bb.num_passes(cfg))]
    assert "_ld" in names[0] and all("_ld" not in n for n in names[1:])
    assert "_st" in names[-1] and all("_st" not in n for n in names[:-1])
    D = cfg.dim
    n_in = int(np.prod([cfg.shape[d] for d in range(D + 2)]))
    if t == 2:
        n_in = n_in // cfg.shape[1] * (cfg.shape[1] // 2 + 1)
    x = oracle.uniform_real(12, n_in) if t == 1 else oracle.uniform_complex(12, n_in)
    x = x.astype(np.float32 if t == 1 else np.complex64)
    y = _run_nd_passes(cfg, x)
    y_ref = _run_nd_passes(plain, (x * np.float32(0.125)).astype(x.dtype))
    assert np.array_equal(y, (y_ref * np.float32(4.0)).astype(y_ref.dtype))
