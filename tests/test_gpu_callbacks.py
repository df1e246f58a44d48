"""Load/store callbacks on the GPU (reference: test/callback.cpp), compared bitwise."""
import numpy as np
import pytest

import bbfft_amd as bb
import callback_sources as cbs
import oracle
from gpu_util import execute, extent, have_gpu, io_shapes

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not have_gpu(), reason="needs a HIP GPU")]

CASES = [(fp, M, N) for fp in (4, 8) for M in (1, 32) for N in (8, 64, 212)]


def _real_t(fp):
    return "float" if fp == 4 else "double"


@pytest.mark.parametrize("fp,M,N", CASES)
def test_load_callback_zero_padding(fp, M, N):
    """c2r of 2N with a zero-padding load callback == c2r of explicitly zero-padded data."""
    K = 16
    N_ext = 2 * N
    n_ref, n_have = N_ext // 2 + 1, N // 2 + 1
    ref_istride = [1, M, M * n_ref]
    ostride = [1, M, M * N_ext]
    cfg_ref = bb.configuration(1, [M, N_ext, K], bb.precision(fp), bb.direction.backward,
                               bb.transform_type.c2r, ref_istride, ostride)
    cfg = bb.configuration(1, [M, N_ext, K], bb.precision(fp), bb.direction.backward,
                           bb.transform_type.c2r, ref_istride, ostride,
                           bb.user_module(cbs.load_hip(_real_t(fp), M, n_ref, n_have), "load"))
    X = oracle.uniform_complex(21, M * n_have * K).reshape(K, n_have, M)
    X_ref = np.zeros((K, n_ref, M), dtype=np.complex128)
    X_ref[:, :n_have, :] = X
    x_ref = execute(cfg_ref, X_ref.ravel())
    # the input buffer only holds the compact N/2+1 bins; pad it to the plan's extent
    compact = np.zeros(X_ref.size, dtype=np.complex128)
    compact[:X.size] = X.ravel()
    x = execute(cfg, compact)
    assert np.array_equal(x, x_ref)


@pytest.mark.parametrize("lang", ["hip", "opencl_c"])
@pytest.mark.parametrize("fp,M,N", CASES)
def test_store_callback_truncate_scale(fp, M, N, lang):
    """r2c with a store callback that keeps N/4 bins scaled by 1/N == r2c output * T(1/N)."""
    K = 16
    n_ref, n_cut = N // 2 + 1, N // 4
    istride = [1, M, M * N]
    ref_ostride = [1, M, M * n_ref]
    real_t = _real_t(fp)
    scale = float(np.float32(1.0 / N)) if fp == 4 else 1.0 / N
    src = (cbs.store_hip if lang == "hip" else cbs.store_opencl)(real_t, M, n_ref, n_cut,
                                                                 scale.hex())
    language = bb.kernel_language.hip if lang == "hip" else bb.kernel_language.opencl_c
    cfg_ref = bb.configuration(1, [M, N, K], bb.precision(fp), bb.direction.forward,
                               bb.transform_type.r2c, istride, ref_ostride)
    cfg = bb.configuration(1, [M, N, K], bb.precision(fp), bb.direction.forward,
                           bb.transform_type.r2c, istride, ref_ostride,
                           bb.user_module(src, None, "store", language))
    x = oracle.uniform_real(22, M * N * K)
    X_ref = execute(cfg_ref, x).reshape(K, n_ref, M)
    # the store callback writes a (M, n_cut, K) tensor into the output buffer
    real_np = np.float32 if fp == 4 else np.float64
    X = execute(cfg, x)[:M * n_cut * K].reshape(K, n_cut, M)
    expect_re = (X_ref.real[:, :n_cut, :].astype(real_np) * real_np(scale))
    expect_im = (X_ref.imag[:, :n_cut, :].astype(real_np) * real_np(scale))
    assert np.array_equal(X.real, expect_re)
    assert np.array_equal(X.imag, expect_im)
    # and against the FP64 oracle directly: the first N/4 bins of the r2c, times 1/N
    ref = oracle.fft(1, [M, N, K], istride, ref_ostride, 1, -1, x).reshape(K, n_ref, M)
    ref = ref[:, :n_cut, :] * (1.0 / N)
    assert oracle.rel_l2(X.ravel(), ref.ravel()) <= oracle.tolerance(fp, N)


@pytest.mark.parametrize("desc", ["scfo8.256*4096", "srfo8.256*4096"])
def test_callback_plan_at_tuned_shape_matches_plain_plan(desc):
    """A tuned shape whose table entry carries the load-first tile loop, with an identity load
    callback: bitwise equal to the plain plan."""
    cfg_ref = bb.parse_fft_descriptor(desc)
    cfg = bb.parse_fft_descriptor(desc)
    t = "float" if cfg.type == bb.transform_type.r2c else "float2"
    cfg.callbacks = bb.user_module(f"__device__ {t} ld({t}* p, size_t o) {{ return p[o]; }}",
                                   "ld")
    M, N, K = cfg.shape[:3]
    if cfg.type == bb.transform_type.r2c:
        x = oracle.uniform_real(23, M * N * K)
    else:
        x = oracle.uniform_complex(23, M * N * K)
    assert np.array_equal(execute(cfg, x), execute(cfg_ref, x))


def test_same_callback_name_different_code_in_one_cache():
    """Two store callbacks named alike with different scale constants, built through one
    cache: each plan runs its own code."""
    M, N, K = 1, 64, 16
    x = oracle.uniform_real(24, M * N * K)
    cfg_ref = bb.parse_fft_descriptor(f"srfo{M}.{N}*{K}")
    X = execute(cfg_ref, x)
    cache = bb.Cache()
    outs = []
    for scale in ("2.0f", "3.0f"):
        cfg = bb.parse_fft_descriptor(f"srfo{M}.{N}*{K}")
        cfg.callbacks = bb.user_module(
            "__device__ void store(float2* o, size_t i, float2 v) "
            f"{{ v.x *= {scale}; v.y *= {scale}; o[i] = v; }}", None, "store")
        outs.append(execute(cfg, x, plan=bb.make_plan(cfg, cache=cache)))
    for out, scale in zip(outs, (np.float32(2), np.float32(3))):
        assert np.array_equal(out.real, X.real * scale)
        assert np.array_equal(out.imag, X.imag * scale)


@pytest.mark.parametrize("fp,M,N", [(4, 1, 64), (4, 32, 212), (8, 1, 212), (8, 32, 8)])
def test_store_callback_truncate_scale_against_oracle(fp, M, N):
    """The truncate-and-scale store callback against the FP64 oracle's r2c times 1/N."""
    K = 16
    n_ref, n_cut = N // 2 + 1, N // 4
    istride = [1, M, M * N]
    ostride = [1, M, M * n_ref]
    real_t = _real_t(fp)
    scale = float(np.float32(1.0 / N)) if fp == 4 else 1.0 / N
    cfg = bb.configuration(1, [M, N, K], bb.precision(fp), bb.direction.forward,
                           bb.transform_type.r2c, istride, ostride,
                           bb.user_module(cbs.store_hip(real_t, M, n_ref, n_cut, scale.hex()),
                                          None, "store"))
    x = oracle.uniform_real(25, M * N * K)
    xr = x.astype(np.float32).astype(np.float64) if fp == 4 else x
    ref = oracle.fft(1, [M, N, K], istride, ostride, 1, -1, xr).reshape(K, n_ref, M)
    got = execute(cfg, x)[:M * n_cut * K].reshape(K, n_cut, M)
    assert oracle.rel_l2(got, ref[:, :n_cut, :] * scale) <= oracle.tolerance(fp, N)


_ND_SCALE_SRC = """
__device__ {t2} ld({t2}* in, size_t offset) {{
    {t2} v = in[offset]; v.x *= ({t})0.125; v.y *= ({t})0.125; return v;
}}
__device__ {t} ldr({t}* in, size_t offset) {{ return in[offset] * ({t})0.125; }}
__device__ void st({t2}* out, size_t offset, {t2} v) {{
    v.x *= ({t})4; v.y *= ({t})4; out[offset] = v;
}}
__device__ void str_({t}* out, size_t offset, {t} v) {{ out[offset] = v * ({t})4; }}
"""


def _nd_callbacks(cfg, load=True, store=True):
    t = int(cfg.type)
    src = _ND_SCALE_SRC.format(t=_real_t(int(cfg.fp)), t2=_real_t(int(cfg.fp)) + "2")
    return bb.user_module(src, ("ldr" if t == 1 else "ld") if load else None,
                          ("str_" if t == 2 else "st") if store else None,
                          bb.kernel_language.hip)


def _nd_input(cfg, seed):
    t = int(cfg.type)
    n = int(np.prod([cfg.shape[d] for d in range(cfg.dim + 2)]))
    if t == 2:
        n = n // cfg.shape[1] * (cfg.shape[1] // 2 + 1)
    return oracle.uniform_real(seed, n) if t == 1 else oracle.uniform_complex(seed, n)


@pytest.mark.parametrize("desc,load,store,fused", [
    ("scfo16x8x4*3", True, True, False), ("srfo16x8x4*3", True, True, False),
    ("srbo16x8x4*3", True, True, False), ("dcbo2.12x10*5", True, True, False),
    # the callback outside the fused pair of passes (r2c: last pass, c2r: first pass)
    ("srfo64x64x64*2", False, True, True), ("srbo64x64x64*2", True, False, True)])
def test_nd_callbacks(desc, load, store, fused, monkeypatch):
    """Callbacks on 2-D/3-D plans (extension): load in the first pass, store in the last, each
    with the plan's own input / output offsets; the passes without callbacks still fuse.
    Scaling by powers of two is exact: the result equals the plain plan's (with the same
    passes, fused or not) on input / 8, times 4."""
    if not fused:
        monkeypatch.setenv("BBFFT_FUSE", "off")  # the reference plan's passes, unfused too
    cfg = bb.parse_fft_descriptor(desc)
    plain = bb.parse_fft_descriptor(desc)
    cfg.callbacks = _nd_callbacks(cfg, load, store)
    x = _nd_input(cfg, 31)
    y = execute(cfg, x)
    y_ref = execute(plain, x * (0.125 if load else 1.0))
    assert np.array_equal(y, y_ref * (4 if store else 1))
    if fused:
        p = bb.make_plan(cfg, cache=bb.default_cache())
        assert any("fused2" in l["kernel"] for l in p.launches())


@pytest.mark.parametrize("desc,load,store,ok", [
    ("srbo16x8x4*3", False, True, True),    # adjoint: r2c-type, store in its last (c2c) pass
    ("srfo16x8x4*3", True, False, True),    # adjoint: c2r-type, load in its first (c2c) pass
    ("scfo16x8x4*3", True, True, True),
    ("srbo16x8x4*3", True, False, False),   # load would run in the adjoint's real pass
    ("srfo1.16*3", False, True, False),     # 1-D real adjoint
])
def test_adjoint_callbacks(desc, load, store, ok, monkeypatch):
    """Adjoint plans take the configuration's callbacks on their own input and output where
    those run in a c2c pass; elsewhere plan creation is refused."""
    monkeypatch.setenv("BBFFT_FUSE", "off")  # same passes with and without callbacks
    cfg = bb.parse_fft_descriptor(desc)
    plain = bb.parse_fft_descriptor(desc)
    cfg.callbacks = _nd_callbacks(bb.adjoint_configuration(plain), load, store)
    if not ok:
        with pytest.raises(bb.BadConfiguration):
            bb.make_adjoint_plan(cfg, cache=bb.default_cache())
        return
    adj = bb.adjoint_configuration(plain)
    x = _nd_input(adj, 32)
    p = bb.make_adjoint_plan(cfg, cache=bb.default_cache())
    p_ref = bb.make_adjoint_plan(plain, cache=bb.default_cache())
    y = execute(adj, x, plan=p)
    y_ref = execute(adj, x * (0.125 if load else 1.0), plan=p_ref)
    assert np.array_equal(y, y_ref * (4 if store else 1))


@pytest.mark.parametrize("desc", ["scfi16x8x4*3", "srfi16x8x4*3", "srbi16x8x4*3"])
def test_nd_callbacks_in_place(desc, monkeypatch):
    """In-place 2-D/3-D plans with callbacks: the first pass reads the buffer it (or a later
    pass) overwrites; same bits as the plain in-place plan on scaled data."""
    monkeypatch.setenv("BBFFT_FUSE", "off")
    cfg = bb.parse_fft_descriptor(desc)
    plain = bb.parse_fft_descriptor(desc)
    cfg.callbacks = _nd_callbacks(cfg)
    ishape, _ = io_shapes(cfg)
    n = extent(cfg.dim, ishape, cfg.istride)
    t = int(cfg.type)
    x = oracle.uniform_real(33, n) if t == 1 else oracle.uniform_complex(33, n)
    y = execute(cfg, x, inplace=True)
    y_ref = execute(plain, x * 0.125, inplace=True)
    if t == 2:  # real rows padded to 2 (N1/2 + 1): the padding is not output
        n1, row = cfg.shape[1], 2 * (cfg.shape[1] // 2 + 1)
        rows = (y.size + row - 1) // row
        idx = (np.arange(rows)[:, None] * row + np.arange(n1)[None, :]).ravel()
        y, y_ref = y[idx], y_ref[idx]
    assert np.array_equal(y, y_ref * 4)
