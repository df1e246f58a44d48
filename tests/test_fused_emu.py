"""Fused passes of 2-D/3-D plans on the CPU (fiber emulator): the plane kernel (passes over
modes 1 and 2 with the plane in LDS, sharing it with the stage scratch) chained with the
remaining pass exactly as the plan chains them, against the FP64 oracle."""
import numpy as np
import pytest

import bbfft_amd as bb
import oracle
from emu import runner


def _cfg(M, dims, K, fp, t, inplace=False):
    d = 1 if t == 2 else -1
    cfg = bb.configuration(len(dims), [M] + list(dims) + [K], bb.precision(fp),
                           bb.direction(d), bb.transform_type(t))
    cfg.set_strides_default(inplace)
    return cfg


def _run_plan(cfg, x, n_out):
    """The plan's steps: the fused pair replaces two passes; other passes as usual."""
    fc = bb.fused_configuration(cfg)
    assert fc is not None
    D = cfg.dim
    fp = int(cfg.fp)
    t = int(cfg.type)
    ct = np.complex64 if fp == 4 else np.complex128
    rt = np.float32 if fp == 4 else np.float64
    M, K = cfg.shape[0], cfg.shape[D + 1]
    n_c = [cfg.shape[1] // 2 + 1 if t != 0 else cfg.shape[1]] + list(cfg.shape[2:D + 1])
    inter = M * int(np.prod(n_c)) * K  # packed complex intermediate
    first = fc["first_pass"]
    if D == 2:
        out = np.zeros(n_out, rt if t == 2 else ct)
        return runner.run_fused(bb, cfg, x, out, grid_cap=3)
    if first == 0:  # fused modes 1, 2 then the pass over mode 3
        mid = np.zeros(inter, ct)
        runner.run_fused(bb, cfg, np.ascontiguousarray(x), mid, grid_cap=5)
        return runner.run(bb, cfg, mid, n_out, pass_index=2)
    # c2r: the pass over mode 3 first, then modes 2, 1 fused
    mid = runner.run(bb, cfg, x, inter, pass_index=0)
    out = np.zeros(n_out, rt)
    return runner.run_fused(bb, cfg, mid, out, grid_cap=5)


@pytest.mark.parametrize("fp", [4, 8])
@pytest.mark.parametrize("t", [0, 1, 2])
@pytest.mark.parametrize("M,dims,K", [(1, [16, 12], 3), (2, [8, 6, 5], 2), (1, [10, 8, 4], 3),
                                      (3, [6, 9], 2), (1, [15, 7, 4], 2)])
def test_fused_vs_oracle(fp, t, M, dims, K):
    cfg = _cfg(M, dims, K, fp, t)
    shape = cfg.shape
    D = len(dims)
    half = list(shape)
    half[1] = dims[0] // 2 + 1
    ishape, oshape = (shape, half) if t == 1 else (half, shape) if t == 2 else (shape, shape)
    n_in = oracle.extent(D, ishape, cfg.istride)
    n_out = oracle.extent(D, oshape, cfg.ostride)
    x = oracle.uniform_real(5, n_in) if t == 1 else oracle.uniform_complex(5, n_in)
    if fp == 4:
        x = x.astype(np.float32 if t == 1 else np.complex64)
    ref = oracle.fft(D, shape, cfg.istride, cfg.ostride, t, int(cfg.dir), x, out_size=n_out)
    y = _run_plan(cfg, x, n_out)
    assert oracle.rel_l2(y, ref) <= oracle.tolerance(fp, int(np.prod(dims)))


@pytest.mark.parametrize("t,M,dims,K", [(0, 1, [16, 12], 5), (1, 1, [16, 12], 5),
                                        (2, 1, [16, 12], 5), (1, 2, [8, 6, 5], 3),
                                        (2, 1, [10, 8, 4], 3), (0, 3, [6, 9], 4)])
def test_fused_staged_matches_unstaged(t, M, dims, K, monkeypatch):
    """The staged kernel (next plane's input streamed into LDS during the current plane) gives
    the same bits as the kernel whose first pass reads global memory itself, with work-groups
    walking several planes (grid smaller than the plane count)."""
    cfg = _cfg(M, dims, K, 4, t)
    shape = cfg.shape
    D = len(dims)
    half = list(shape)
    half[1] = dims[0] // 2 + 1
    ishape = half if t == 2 else shape
    n_in = oracle.extent(D, ishape, cfg.istride)
    x = oracle.uniform_real(9, n_in) if t == 1 else oracle.uniform_complex(9, n_in)
    x = x.astype(np.float32 if t == 1 else np.complex64)
    ys = []
    for stage in ("on", "off"):
        if stage == "off":
            monkeypatch.setenv("BBFFT_FUSE_STAGE", "off")
        fc = bb.fused_configuration(cfg)
        src = bb.generate_fused_source(cfg)
        assert ("_stage[" in src) == (stage == "on")
        oshape = half if t == 1 else shape
        n_out = oracle.extent(D, oshape, cfg.ostride)
        ys.append(_run_plan(cfg, x, n_out))
    assert np.array_equal(ys[0], ys[1])


# every variant setting off; VARIANTS override it (an empty dict: the plan's defaults)
PLAIN = {"BBFFT_FUSE_DMA": "4", "BBFFT_FUSE_SYNC": "full", "BBFFT_FUSE_TABLES": "const",
         "BBFFT_FUSE_EARLY": "off", "BBFFT_FUSE_DMAWAVES": "all", "BBFFT_FUSE_SWIZZLE": "off",
         "BBFFT_FUSE_TABPAD": "off"}
VARIANTS = [
    None,
    {"BBFFT_FUSE_DMA": "16", "BBFFT_FUSE_SYNC": "lds", "BBFFT_FUSE_TABLES": "lds"},
    {"BBFFT_FUSE_DMA": "16"},
    {"BBFFT_FUSE_TABLES": "lds"},
    {"BBFFT_FUSE_STAGE": "off", "BBFFT_FUSE_SYNC": "lds", "BBFFT_FUSE_TABLES": "lds",
     "BBFFT_FUSE_WAVES": "8"},
    {"BBFFT_FUSE_DMA": "16", "BBFFT_FUSE_SYNC": "lds", "BBFFT_FUSE_TABLES": "lds",
     "BBFFT_FUSE_DMAWAVES": "split"},
    {"BBFFT_FUSE_DMA": "16", "BBFFT_FUSE_SYNC": "lds", "BBFFT_FUSE_TABLES": "lds",
     "BBFFT_FUSE_EARLY": "on"},
    {"BBFFT_FUSE_SYNC": "lds", "BBFFT_FUSE_TABLES": "lds", "BBFFT_FUSE_EARLY": "on",
     "BBFFT_FUSE_DMAWAVES": "split"},
    {"BBFFT_FUSE_TABLES": "lds", "BBFFT_FUSE_TABPAD": "on"},
]


@pytest.mark.parametrize("variant", range(len(VARIANTS)))
@pytest.mark.parametrize("t,M,dims,K", [(1, 1, [32, 24], 5), (2, 1, [32, 24], 5),
                                        (0, 2, [16, 12, 5], 3), (2, 1, [32, 16, 4], 3),
                                        (1, 1, [16, 64], 9), (0, 1, [8, 32, 3], 5)])
def test_fused_variants_match_default(t, M, dims, K, variant, monkeypatch):
    """The fused kernel's variants (16-byte staging loads, barriers ordering LDS only,
    twiddle tables in LDS, occupancy target) compute the same bits as the default kernel."""
    cfg = _cfg(M, dims, K, 4, t)
    shape = cfg.shape
    D = len(dims)
    half = list(shape)
    half[1] = dims[0] // 2 + 1
    ishape = half if t == 2 else shape
    oshape = half if t == 1 else shape
    n_in = oracle.extent(D, ishape, cfg.istride)
    n_out = oracle.extent(D, oshape, cfg.ostride)
    x = oracle.uniform_real(4, n_in) if t == 1 else oracle.uniform_complex(4, n_in)
    x = x.astype(np.float32 if t == 1 else np.complex64)
    for k, v in PLAIN.items():
        monkeypatch.setenv(k, v)
    base_name = bb.fused_configuration(cfg)["name"]
    y0 = _run_plan(cfg, x, n_out)
    var = VARIANTS[variant]
    if var is None:  # the defaults
        for k in PLAIN:
            monkeypatch.delenv(k)
        var = {}
    for k, v in var.items():
        monkeypatch.setenv(k, v)
    fc = bb.fused_configuration(cfg)
    src = bb.generate_fused_source(cfg)
    if VARIANTS[variant] is not None or t == 1:
        assert fc["name"] != base_name
    if var.get("BBFFT_FUSE_DMA") == "16":
        assert "BBFFT_LDS_DMA16(" in src
    if var.get("BBFFT_FUSE_TABLES") == "lds":
        assert "_tables(threadIdx.x" in src
    if var.get("BBFFT_FUSE_TABPAD") == "on" and dims != [16, 12, 5]:
        assert "_l[(q / " in src  # table rows padded
    if var.get("BBFFT_FUSE_DMAWAVES") == "split" and dims in ([16, 64], [8, 32, 3]):
        assert "copying waves\n" in src
    if var.get("BBFFT_FUSE_EARLY") == "on":
        assert "(next_input, stage);" in src
    y1 = _run_plan(cfg, x, n_out)
    assert np.array_equal(y0, y1)


@pytest.mark.parametrize("t,M,dims,K", [(1, 1, [128, 16], 3), (0, 1, [64, 16], 3),
                                        (1, 1, [256, 8], 2)])
def test_fused_swizzled_stage_matches_plain(t, M, dims, K, monkeypatch):
    """Rows of 512 bytes (or a multiple) sit in the staging buffer with their 64-byte segments
    permuted row by row (the 8 rows a wave reads at once then use different LDS banks); the
    kernel gives the same bits as with the plain layout and as the unstaged kernel."""
    cfg = _cfg(M, dims, K, 4, t)
    shape = cfg.shape
    D = len(dims)
    half = list(shape)
    half[1] = dims[0] // 2 + 1
    oshape = half if t == 1 else shape
    n_in = oracle.extent(D, shape, cfg.istride)
    n_out = oracle.extent(D, oshape, cfg.ostride)
    x = oracle.uniform_real(6, n_in) if t == 1 else oracle.uniform_complex(6, n_in)
    x = x.astype(np.float32 if t == 1 else np.complex64)
    ys = []
    for env in ({"BBFFT_FUSE_SWIZZLE": "on"}, {"BBFFT_FUSE_SWIZZLE": "off"},
                {"BBFFT_FUSE_STAGE": "off"}, {}):
        for k in ("BBFFT_FUSE_SWIZZLE", "BBFFT_FUSE_STAGE"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        src = bb.generate_fused_source(cfg)
        # on by default for r2c first passes
        swz = env.get("BBFFT_FUSE_SWIZZLE") == "on" or (not env and t == 1)
        assert ("rows permuted" in src) == swz
        ys.append(_run_plan(cfg, x, n_out))
    assert all(np.array_equal(ys[0], y) for y in ys[1:])
    ref = oracle.fft(D, shape, cfg.istride, cfg.ostride, t, int(cfg.dir),
                     x.astype(np.float64 if t == 1 else np.complex128), out_size=n_out)
    assert oracle.rel_l2(ys[0], ref) <= oracle.tolerance(4, int(np.prod(dims)))


def test_fused_copy_by_waves_idle_in_first_pass(monkeypatch):
    """c2r plane whose first (c2c) pass needs fewer waves than its second: with
    BBFFT_FUSE_DMAWAVES=idle_a the next plane's staging loads are issued right after the first
    pass's first stage by the waves that have no work in the rest of that pass. Same bits as
    the plain kernel and as the unstaged one."""
    cfg = _cfg(1, [32, 64], 3, 4, 2)
    half = list(cfg.shape)
    half[1] = 32 // 2 + 1
    n_in = oracle.extent(2, half, cfg.istride)
    n_out = oracle.extent(2, cfg.shape, cfg.ostride)
    x = oracle.uniform_complex(8, n_in).astype(np.complex64)
    monkeypatch.setenv("BBFFT_FUSE_WGMAX", "256")  # pass a: 136 work-items, pass b: 256
    ys = []
    for env in (PLAIN, {"BBFFT_FUSE_SYNC": "lds", "BBFFT_FUSE_TABLES": "lds",
                        "BBFFT_FUSE_EARLY": "on", "BBFFT_FUSE_DMAWAVES": "idle_a"},
                {"BBFFT_FUSE_STAGE": "off"}):
        for k in list(PLAIN) + ["BBFFT_FUSE_STAGE"]:
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        src = bb.generate_fused_source(cfg)
        if env.get("BBFFT_FUSE_DMAWAVES") == "idle_a":
            assert "threadIdx.x / 64u - 3u < 1u) BBFFT_WAIT_ALL(); // copying waves" in src
            assert "(next_input, stage);" in src
        ys.append(_run_plan(cfg, x, n_out))
    assert all(np.array_equal(ys[0], y) for y in ys[1:])


def test_fused_in_place_layout_r2c():
    """In-place r2c layout (padded real rows) read by the fused kernel."""
    cfg = _cfg(1, [16, 8, 4], 2, 4, 1, inplace=True)
    n_in = oracle.extent(3, cfg.shape, cfg.istride)
    half = list(cfg.shape)
    half[1] = 9
    n_out = oracle.extent(3, half, cfg.ostride)
    x = oracle.uniform_real(6, n_in).astype(np.float32)
    ref = oracle.fft(3, cfg.shape, cfg.istride, cfg.ostride, 1, -1, x, out_size=n_out)
    y = _run_plan(cfg, x, n_out)
    m, n1, n2, n3, k = np.meshgrid(range(1), range(9), range(8), range(4), range(2), indexing="ij")
    sel = (n1 * cfg.ostride[1] + n2 * cfg.ostride[2] + n3 * cfg.ostride[3] +
           k * cfg.ostride[4]).ravel()
    assert oracle.rel_l2(y[sel], ref[sel]) <= oracle.tolerance(4, 512)


def test_staging_cache_policy_follows_the_other_passes(monkeypatch):
    """The plane kernel's staging loads take the non-temporal cache policy when every other
    pass of the plan stores plainly (the 32 x 128^3 real plans, whose third pass has a plain
    store entry in the tuning table), and plain loads when another pass streams its stores or
    the plan has no other pass; BBFFT_FUSE_DMA_NT overrides."""
    def nt(desc):
        return "BBFFT_LDS_DMA16_NT(src" in bb.generate_fused_source(bb.parse_fft_descriptor(desc))
    assert nt("srfo128x128x128*32") and nt("srbo128x128x128*32")
    assert not nt("scfo64x128x128*32") and not nt("srfo64x64x64*256")
    assert not nt("srfo128x128*2048") and not nt("dcfo64x64*1024")
    monkeypatch.setenv("BBFFT_FUSE_DMA_NT", "on")
    assert nt("srfo128x128*2048")
    monkeypatch.setenv("BBFFT_FUSE_DMA_NT", "off")
    assert not nt("srfo128x128x128*32")


def test_fusion_policy(monkeypatch):
    """Planes of at most 96 KB of LDS fuse, among them the 128 x 65 fp32 plane of the
    32 x 128^3 real transforms (70 KB); larger planes (the 138 KB fp32 128 x 128 plane) are left
    to the unfused passes unless BBFFT_FUSE=on; planes that do not fit never fuse; small
    problems fuse (launch-bound)."""
    c3 = bb.parse_fft_descriptor("srfo128x128x128*32")
    c2r = _cfg(1, [128, 128, 128], 32, 4, 2)
    assert bb.fused_configuration(c3) is not None
    assert bb.fused_configuration(c2r) is not None
    big = bb.parse_fft_descriptor("scfo128x128*512")
    assert bb.fused_configuration(big) is None
    assert bb.fused_configuration(bb.parse_fft_descriptor("scfo64x64x64*16")) is not None
    assert bb.fused_configuration(bb.parse_fft_descriptor("scfo256x256*4")) is None
    assert bb.fused_configuration(bb.parse_fft_descriptor("scfo64x64*100")) is not None
    monkeypatch.setenv("BBFFT_FUSE", "on")
    fc = bb.fused_configuration(c3)
    assert fc["first_pass"] == 0 and fc["planes"] == 128 * 32 and fc["plane_elems"] == 65 * 128
    # plane + stage scratch <= 80 KB, plus the 64 KB staging buffer of the next plane's input
    assert fc["lds_bytes"] <= 160 * 1024 and fc["group_size"] <= 1024
    monkeypatch.setenv("BBFFT_FUSE_STAGE", "off")
    unstaged = bb.fused_configuration(c3)
    assert unstaged["lds_bytes"] <= 80 * 1024 and unstaged["name"] != fc["name"]
    assert fc["lds_bytes"] == unstaged["lds_bytes"] + 4 * 128 * 128
    monkeypatch.delenv("BBFFT_FUSE_STAGE")
    fr = bb.fused_configuration(c2r)
    assert fr["first_pass"] == 1 and fr["kb"] == 128 and fr["ka"] == 1
    assert bb.fused_configuration(bb.parse_fft_descriptor("scfo256x256*4")) is None
    assert bb.fused_configuration(big)["lds_bytes"] > 96 * 1024
    monkeypatch.setenv("BBFFT_FUSE", "off")
    assert bb.fused_configuration(bb.parse_fft_descriptor("scfo64x64*100")) is None


def test_macro_definitions_agree_across_kernels():
    """Ahead-of-time archives compile many kernels in one translation unit, where the first
    definition of an include-guarded macro wins: every macro that generated sources define must
    therefore be defined identically by all of them (kernel-specific behaviour never goes
    through a macro)."""
    import re
    srcs = [bb.generate_fused_source(bb.parse_fft_descriptor(d))
            for d in ("srfo128x128x128*32", "srbo128x128x128*32", "scfo64x64x64*8",
                      "srfo1.16x64*300")]
    for d in ("scfo8.256*32768", "srfo3.128*33", "srbo5.96*33", "dcbo8.60*5", "scfo1.509*33"):
        cfg = bb.parse_fft_descriptor(d)
        srcs.append(bb.generate_source(cfg, 0, bb.GFX950_INFO))
    defs = {}
    for src in srcs:
        for m in re.finditer(r"^#define\s+(\w+)(.*)$", src, re.M):
            defs.setdefault(m.group(1), set()).add(m.group(2).strip())
    assert "BBFFT_LDS_SYNC" in defs
    for name, bodies in defs.items():
        assert len(bodies) == 1, (name, bodies)
