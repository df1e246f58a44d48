"""GPU parity of real transforms (reference: test/r2c.cpp) through the C ABI."""
import numpy as np
import pytest

import bbfft_amd as bb
import gpu_cases
import oracle
from gpu_util import approx_equal, execute, grid, have_gpu, io_shapes, offsets, ref_tol

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not have_gpu(), reason="needs a HIP GPU")]

TAU = 2.0 * np.pi


def _split(cfg):
    D = cfg.dim
    return cfg.shape[0], list(cfg.shape[1:D + 1]), cfg.shape[D + 1]


def _half_delta(nd, b, N):
    return (((nd - b) % N == 0).astype(float) + ((nd + b) % N == 0).astype(float)) / 2.0


@pytest.mark.parametrize("desc", gpu_cases.r2c_1d_oop() + gpu_cases.r2c_1d_inplace()
                         + gpu_cases.r2c_nd())
def test_r2c_forward_cosine(desc):
    """x = s(k) prod_d cos(2 pi b_d n_d / N_d) / N_d -> X = s(k) prod_d half-deltas at +-b_d
    (test/r2c.cpp:40-56,82-105,143-169)."""
    cfg = bb.parse_fft_descriptor(desc)
    M, Ns, K = _split(cfg)
    ishape, oshape = io_shapes(cfg)
    idx = grid(ishape)
    m, k = idx[0], idx[-1]
    val = 1.0 + k / K
    for d, N in enumerate(Ns):
        b = (m + k) % N
        val = val * np.cos(TAU * ((b * idx[d + 1]) % N) / N) / N
    ioff = offsets(idx, cfg.istride[:cfg.dim + 2])
    x = np.zeros(ioff.max() + 1)
    x[ioff] = val
    inplace = desc[3] == "i"
    y = execute(cfg, x, inplace=inplace)
    odx = grid(oshape)
    m, k = odx[0], odx[-1]
    ref = 1.0 + k / K
    for d, N in enumerate(Ns):
        ref = ref * _half_delta(odx[d + 1], (m + k) % N, N)
    out = y[offsets(odx, cfg.ostride[:cfg.dim + 2])]
    eps = ref_tol(cfg.fp, int(np.prod(Ns)))
    assert approx_equal(out.real, ref, eps).all()
    assert approx_equal(out.imag, 0.0, eps).all()


@pytest.mark.parametrize("desc,pollute", [(d, False) for d in gpu_cases.c2r_1d_oop()
                                          + gpu_cases.c2r_nd()]
                         + [(d, True) for d in gpu_cases.c2r_polluted()])
def test_c2r_backward_cosine(desc, pollute):
    """X = s(k) prod_d half-deltas -> x = s(k) prod_d cos(2 pi b_d n_d / N_d); with `pollute`
    imag(X[0]) = 1 + m + k must be ignored (test/r2c.cpp:58-80,120-141,246-324)."""
    cfg = bb.parse_fft_descriptor(desc)
    M, Ns, K = _split(cfg)
    ishape, oshape = io_shapes(cfg)
    idx = grid(ishape)
    m, k = idx[0], idx[-1]
    val = (1.0 + k / K).astype(np.complex128)
    for d, N in enumerate(Ns):
        val = val * _half_delta(idx[d + 1], (m + k) % N, N)
    if pollute:
        zero = np.all([i == 0 for i in idx[1:-1]], axis=0)
        val = val + 1j * zero * (1.0 + m + k)
    ioff = offsets(idx, cfg.istride[:cfg.dim + 2])
    X = np.zeros(ioff.max() + 1, dtype=np.complex128)
    X[ioff] = val
    inplace = desc[3] == "i"
    y = execute(cfg, X, inplace=inplace)
    odx = grid(oshape)
    m, k = odx[0], odx[-1]
    ref = 1.0 + k / K
    for d, N in enumerate(Ns):
        ref = ref * np.cos(TAU * ((((m + k) % N) * odx[d + 1]) % N) / N)
    out = y[offsets(odx, cfg.ostride[:cfg.dim + 2])]
    eps = ref_tol(cfg.fp, int(np.prod(Ns)))
    assert approx_equal(out, ref, eps).all()


@pytest.mark.parametrize("case", gpu_cases.r2c_identity())
def test_r2c_c2r_identity(case):
    """c2r(r2c(x)) / prod(N) == x (test/r2c.cpp:376-497), seeded."""
    p, place, M, Ns, K = case
    fwd, bwd = gpu_cases.identity_descriptors(case)
    cfg, icfg = bb.parse_fft_descriptor(fwd), bb.parse_fft_descriptor(bwd)
    ishape, _ = io_shapes(cfg)
    idx = grid(ishape)
    ioff = offsets(idx, cfg.istride[:cfg.dim + 2])
    x = np.zeros(ioff.max() + 1)
    x[ioff] = oracle.uniform_real(5, ioff.size)
    inplace = place == "i"
    X = execute(cfg, x, inplace=inplace)
    xr = execute(icfg, X, inplace=inplace) / np.prod(Ns)
    got = xr[offsets(idx, icfg.ostride[:cfg.dim + 2])]
    eps = ref_tol(cfg.fp, int(np.prod(Ns)))
    assert approx_equal(got, x[ioff], eps).all()


# the tuned M=8 real-transform entries (persistent grids, load-first loops), odd K included
_M8_REAL = [f"{p}r{d}o8.{n}*{k}" for p in "sd" for d in "fb"
            for n, k in ((16, 301), (32, 77), (64, 40), (128, 33), (256, 17), (512, 9),
                         (60, 45), (105, 21), (243, 13), (384, 11))]


@pytest.mark.parametrize("desc", ["srfo3.16*5", "drfo3.27*6", "srfi1.64*7", "srfo2.5x6*3",
                                  "drfo1.8x6x5*2", "srbo3.16*5", "drbo3.27*6", "srbi1.64*7",
                                  "srbo2.5x6*3", "drbo1.8x6x5*2", "srbo2.16*3"] + _M8_REAL)
def test_real_random_vs_oracle(desc):
    """Seeded random inputs vs the FP64 oracle (rel-L2 <= 8 eps log2 N). c2r inputs have a
    non-zero imaginary part in the DC bin (ignored) and, for even N_1, in the Nyquist bin
    (the reference adds -imag(X[N/2]) to every sample)."""
    cfg = bb.parse_fft_descriptor(desc)
    M, Ns, K = _split(cfg)
    ishape, oshape = io_shapes(cfg)
    idx = grid(ishape)
    ioff = offsets(idx, cfg.istride[:cfg.dim + 2])
    if int(cfg.type) == 1:
        x = np.zeros(ioff.max() + 1)
        x[ioff] = oracle.uniform_real(9, ioff.size)
    else:
        x = np.zeros(ioff.max() + 1, dtype=np.complex128)
        x[ioff] = oracle.uniform_complex(9, ioff.size)
    y = execute(cfg, x, inplace=desc[3] == "i")
    ref = oracle.fft(cfg.dim, cfg.shape, cfg.istride, cfg.ostride, int(cfg.type),
                     int(cfg.dir), x, out_size=y.size)
    odx = grid(oshape)
    ooff = offsets(odx, cfg.ostride[:cfg.dim + 2])
    assert oracle.rel_l2(y[ooff], ref[ooff]) <= oracle.tolerance(int(cfg.fp), int(np.prod(Ns)))


@pytest.mark.parametrize("desc", [d for d in gpu_cases.c3_descriptors() if d.endswith("*2")])
def test_3d_r2c_c2r_128cubed(desc):
    """128^3 real transforms (2 batches, N(0,1) inputs) vs numpy's pocketfft in float64 (the
    brute-force oracle is too slow at this size; pocketfft is pinned against it in
    tests/test_oracle.py)."""
    cfg = bb.parse_fft_descriptor(desc)
    rng = np.random.default_rng(3)
    xr = rng.standard_normal((2, 128, 128, 128))
    Xr = np.fft.rfftn(xr, axes=(1, 2, 3))
    if int(cfg.type) == 1:
        y = execute(cfg, xr.ravel())
        ref = Xr.ravel()
    else:
        y = execute(cfg, Xr.ravel())
        ref = np.fft.irfftn(Xr, s=(128, 128, 128), axes=(1, 2, 3)).ravel() * 128 ** 3
    assert oracle.rel_l2(y, ref) <= oracle.tolerance(4, 128 ** 3)


@pytest.mark.parametrize("desc", [d for d in gpu_cases.c3_descriptors() if d.endswith("*2")] +
                         ["scfo64x64x64*4", "scbo96x80*40", "drfo48x40x8*3", "drbo48x40x8*3"])
def test_fused_passes_forced(desc, monkeypatch):
    """Plans with the passes over modes 1 and 2 fused (BBFFT_FUSE=on, one plane per
    work-group in LDS) against the unfused plan (BBFFT_FUSE=off) and numpy's pocketfft."""
    cfg = bb.parse_fft_descriptor(desc)
    monkeypatch.setenv("BBFFT_FUSE", "on")
    fc = bb.fused_configuration(cfg)
    assert fc is not None, desc
    fused = bb.make_plan(cfg, cache=bb.default_cache())
    assert any(l["kernel"] == fc["name"] for l in fused.launches())
    monkeypatch.setenv("BBFFT_FUSE", "off")
    plain = bb.make_plan(cfg, cache=bb.default_cache())
    assert all("fused" not in l["kernel"] for l in plain.launches())
    D = cfg.dim
    Ns = list(cfg.shape[1:D + 1])
    B = cfg.shape[D + 1]
    t = int(cfg.type)
    rng = np.random.default_rng(11)
    axes = tuple(range(1, D + 1))
    xr = rng.standard_normal([B] + Ns[::-1])
    if t == 0:
        x = xr + 1j * rng.standard_normal(xr.shape)
        ref = (np.fft.fftn(x, axes=axes) if int(cfg.dir) < 0 else
               np.fft.ifftn(x, axes=axes) * np.prod(Ns)).ravel()
        inp = x.ravel()
    elif t == 1:
        inp, ref = xr.ravel(), np.fft.rfftn(xr, axes=axes).ravel()
    else:
        X = np.fft.rfftn(xr, axes=axes)
        inp, ref = X.ravel(), xr.ravel() * np.prod(Ns)
    y_f = execute(cfg, inp, plan=fused)
    y_p = execute(cfg, inp, plan=plain)
    tol = oracle.tolerance(int(cfg.fp), int(np.prod(Ns)))
    assert oracle.rel_l2(y_f, ref) <= tol
    assert oracle.rel_l2(y_f, y_p) <= tol


@pytest.mark.parametrize("desc", ["srfo128x128x128*4", "srbo128x128x128*4", "scfo64x64x64*8",
                                  "scbo96x80*400", "drfo48x40x8*400"])
def test_fused_staged_vs_unstaged(desc, monkeypatch):
    """Fused kernels that stream the next plane's input into LDS (direct global-to-LDS loads,
    persistent grid; more planes than resident work-groups, so groups walk several planes)
    give the same bits as the fused kernels whose first pass reads global memory itself
    (BBFFT_FUSE_STAGE=off)."""
    cfg = bb.parse_fft_descriptor(desc)
    monkeypatch.setenv("BBFFT_FUSE", "on")
    ys = []
    for stage in ("on", "off"):
        if stage == "off":
            monkeypatch.setenv("BBFFT_FUSE_STAGE", "off")
        plan = bb.make_plan(cfg, cache=bb.default_cache())
        fused = [l for l in plan.launches() if l["kernel"].startswith("bbfft_fused2")]
        assert len(fused) == 1
        if stage == "on":
            fc = bb.fused_configuration(cfg)
            assert fused[0]["grid"] < fc["planes"]
        rng = np.random.default_rng(3)
        ishape, _ = io_shapes(cfg)
        n_in = oracle.extent(cfg.dim, ishape, cfg.istride)
        t = int(cfg.type)
        x = rng.uniform(-1, 1, n_in) if t == 1 else (rng.uniform(-1, 1, n_in) +
                                                      1j * rng.uniform(-1, 1, n_in))
        ys.append(execute(cfg, x, plan=plan))
    assert np.array_equal(ys[0].view(np.uint8), ys[1].view(np.uint8))


# every fused-kernel variant setting off (the plan defaults switch some of them on)
FUSE_PLAIN = {"BBFFT_FUSE_DMA": "4", "BBFFT_FUSE_SYNC": "full", "BBFFT_FUSE_TABLES": "const",
              "BBFFT_FUSE_EARLY": "off", "BBFFT_FUSE_DMAWAVES": "all"}


@pytest.mark.parametrize("desc", ["srfo128x128x128*4", "srbo128x128x128*4", "srfo1.16x64*300",
                                  "scfo64x64x64*8", "drfo48x40x8*400"])
def test_fused_default_variant_vs_plain(desc, monkeypatch):
    """The plan's default fused kernel (for r2c plans: barriers ordering LDS only, LDS twiddle
    tables, 16-byte staging loads issued after pass a's first stage by the waves idle in pass
    b) gives the same bits as the fused kernel with all of these switched off."""
    cfg = bb.parse_fft_descriptor(desc)
    monkeypatch.setenv("BBFFT_FUSE", "on")
    rng = np.random.default_rng(11)
    ishape, _ = io_shapes(cfg)
    n_in = oracle.extent(cfg.dim, ishape, cfg.istride)
    t = int(cfg.type)
    x = rng.uniform(-1, 1, n_in) if t == 1 else (rng.uniform(-1, 1, n_in) +
                                                  1j * rng.uniform(-1, 1, n_in))
    ys = []
    for plain in (True, False):
        for k, v in FUSE_PLAIN.items():
            if plain:
                monkeypatch.setenv(k, v)
            else:
                monkeypatch.delenv(k, raising=False)
        plan = bb.make_plan(cfg, cache=bb.default_cache())
        assert any(l["kernel"].startswith("bbfft_fused2") for l in plan.launches())
        ys.append(execute(cfg, x, plan=plan))
    assert np.array_equal(ys[0].view(np.uint8), ys[1].view(np.uint8))
