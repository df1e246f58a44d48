"""Race checking of the generated kernels on the CPU: each kernel runs once in the fiber
emulator built with ThreadSanitizer (tests/emu/hip_emu.hpp), where every work-item is a fiber
of its own and only the work-group's barriers order their accesses. A report means two
work-items of one group touched the same LDS or global address, at least one writing, with no
barrier between them (in-place transforms included, where input and output alias). Planted
mistakes (a barrier removed from the generated source) show that the check fires."""
import numpy as np
import pytest

import bbfft_amd as bb
from bbfft_amd import ragged
from emu import runner

pytestmark = pytest.mark.skipif(not runner.have_tsan(), reason="no ThreadSanitizer runtime")


def _plain(desc, tuning=None, grid_cap=3):
    """(source, call, in bytes, out bytes, grid, block, in place) of pass 0 of a 1-D plan."""
    cfg = bb.parse_fft_descriptor(desc)
    info = bb.GFX950_INFO
    kc = bb.kernel_configuration(cfg, 0, info, tuning=tuning)
    src = bb.generate_source(cfg, 0, info, tuning=tuning)
    f32 = int(cfg.fp) == 4
    rt, ct = ("float", "float2") if f32 else ("double", "double2")
    t = int(cfg.type)
    in_t = rt if t == 1 else ct
    out_t = rt if t == 2 else ct
    M, N, K = cfg.shape[:3]
    rb = 4 if f32 else 8
    half = N // 2 + 1
    ishape = [M, half if t == 2 else N, K]
    oshape = [M, half if t == 1 else N, K]
    n_in = sum((ishape[i] - 1) * cfg.istride[i] for i in range(3)) + 1
    n_out = sum((oshape[i] - 1) * cfg.ostride[i] for i in range(3)) + 1
    in_b = n_in * (rb if t == 1 else 2 * rb)
    out_b = n_out * (rb if t == 2 else 2 * rb)
    pairs = t != 0 and N % 2 == 1
    kcols = (K + 1) // 2 if pairs else K
    ntiles = -(-M // kc["Mb"]) * (-(-kcols // kc["Kb"]))
    inplace = desc[3] == "i"
    call = f'{kc["name"]}((const {in_t}*)in, ({out_t}*)out, {K}ul)'
    return src, call, max(in_b, out_b) if inplace else in_b, out_b, \
        min(ntiles, grid_cap), kc["work_group_size"], inplace


PLAIN = [
    "scfo8.256*4",     # three stages, 256 work-items, M tile 8
    "scbo3.60*7",      # mixed radices, backward
    "dcfo2.105*3",     # fp64, radix 7 and 15
    "srfo3.128*5",     # r2c, even N: half-length transform and post-pass in LDS
    "srfo2.27*5",      # r2c, odd N: two K columns per transform
    "srbo5.96*5",      # c2r, even N: pre-pass
    "srbo3.25*4",      # c2r, odd N
    "scfi16.63*3",     # in place: input and output alias
    "srfi3.110*4",     # r2c in place (padded rows)
    "scfo1.67*5",      # direct DFT stage (prime above the register limit)
    "dcbo2.254*3",     # direct stage after a radix-2 stage
]


@pytest.mark.parametrize("desc", PLAIN)
def test_plain_kernels_are_race_free(desc):
    ok, report = runner.race_check(*_plain(desc)[:6], in_place=_plain(desc)[6])
    assert ok, report[-3000:]


def test_tile_loop_with_prefetch_is_race_free():
    """Load-first tile loop on a persistent grid: the next tile's inputs are loaded while the
    current tile's stages exchange through LDS."""
    tuning = {"flags": bb.PREFETCH | 3, "groups_per_cu": 1}
    src, call, ib, ob, grid, block, inplace = _plain("scfo8.256*12", tuning=tuning, grid_cap=2)
    assert "tile_next" in src
    ok, report = runner.race_check(src, call, ib, ob, grid, block, in_place=inplace)
    assert ok, report[-3000:]


def _fused(M, dims, K, t, env=None):
    d = 1 if t == 2 else -1
    cfg = bb.configuration(len(dims), [M] + list(dims) + [K], bb.precision.f32,
                           bb.direction(d), bb.transform_type(t))
    cfg.set_strides_default(False)
    fc = bb.fused_configuration(cfg)
    src = bb.generate_fused_source(cfg)
    first = fc["first_pass"]
    in_t = "float" if (t == 1 and first == 0) else "float2"
    out_t = "float" if t == 2 else "float2"
    call = f'{fc["name"]}((const {in_t}*)in, ({out_t}*)out, {fc["planes"]}ul)'
    # input and output of the fused pair: whole planes
    in_b = fc["planes"] * fc["in_plane_stride"] * (4 if in_t == "float" else 8)
    out_b = fc["planes"] * fc["out_plane_stride"] * (4 if out_t == "float" else 8)
    return src, call, in_b, out_b, min(fc["planes"], 3), fc["group_size"]


FUSED_VARIANTS = {
    "default": {},
    "plain": {"BBFFT_FUSE_DMA": "4", "BBFFT_FUSE_SYNC": "full", "BBFFT_FUSE_TABLES": "const",
              "BBFFT_FUSE_EARLY": "off", "BBFFT_FUSE_DMAWAVES": "all", "BBFFT_FUSE_SWIZZLE": "off"},
    "unstaged": {"BBFFT_FUSE_STAGE": "off"},
    "padded_tables": {"BBFFT_FUSE_TABLES": "lds", "BBFFT_FUSE_TABPAD": "on"},
}


@pytest.mark.parametrize("variant", sorted(FUSED_VARIANTS))
@pytest.mark.parametrize("t,M,dims,K", [(1, 1, [16, 64], 9), (2, 1, [32, 16, 4], 3),
                                        (0, 1, [8, 32, 3], 5), (1, 1, [32, 24], 5),
                                        (1, 1, [128, 16], 3)])
def test_fused_kernels_are_race_free(t, M, dims, K, variant, monkeypatch):
    """Fused plane kernels, several planes per work-group: the staged input of the next plane
    lands in LDS while the current plane's second pass runs; the plane shares the LDS with
    both passes' stage scratch."""
    monkeypatch.setenv("BBFFT_FUSE", "on")
    for k, v in FUSED_VARIANTS[variant].items():
        monkeypatch.setenv(k, v)
    ok, report = runner.race_check(*_fused(M, dims, K, t))
    assert ok, report[-3000:]


def test_fused_copy_by_waves_idle_in_first_pass_is_race_free(monkeypatch):
    """The staging copy issued by the waves idle in the first pass (c2r plane, 136 of 256
    work-items busy in pass a) lands while the other waves finish that pass."""
    monkeypatch.setenv("BBFFT_FUSE", "on")
    for k, v in {"BBFFT_FUSE_WGMAX": "256", "BBFFT_FUSE_SYNC": "lds", "BBFFT_FUSE_TABLES": "lds",
                 "BBFFT_FUSE_EARLY": "on", "BBFFT_FUSE_DMAWAVES": "idle_a"}.items():
        monkeypatch.setenv(k, v)
    args = _fused(1, [32, 64], 3, 2)
    assert "copying waves" in args[0]
    ok, report = runner.race_check(*args)
    assert ok, report[-3000:]


def test_ragged_kernel_is_race_free():
    rng = np.random.default_rng(5)
    jobs = ragged.pack([ragged.Job(M=int(rng.integers(1, 4)), N=int(n), K=int(rng.integers(1, 6)),
                                   direction=int(rng.choice([-1, 1])))
                        for n in (12, 60, 100, 128, 243, 16, 64)])
    ok, report = runner.race_check(*_ragged(jobs))
    assert ok, report[-3000:]


def _ragged(jobs, grid=4):
    kc = ragged.kernel_configuration(jobs, 4)
    src = ragged.generate_source(jobs, 4)
    table = ragged.job_table(jobs, 4, grid=grid)
    n = ragged.total_elements(jobs)
    call = (f'{kc["name"]}((const float2*)in, (float2*)out, table, {len(jobs)}ul)')
    return src, call, n * 8, n * 8, grid, kc["group_size"], False, table


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_half_input_mfma_kernel_is_race_free(storage):
    """MFMA split stage: each wave's lanes exchange the matrix operands through the
    emulation's per-wave arrays between two barriers; D lands in LDS for the vector stages."""
    cfg = bb.parse_fft_descriptor("scfo4.512*2")
    flags = bb.half_input_flags(storage, True)
    tuning = {"flags": flags}
    kc = bb.kernel_configuration(cfg, 0, bb.GFX950_INFO, tuning=tuning)
    src = bb.generate_source(cfg, 0, bb.GFX950_INFO, tuning=tuning)
    n = 4 * 512 * 2
    call = f'{kc["name"]}((const unsigned*)in, (float2*)out, 2ul)'
    ok, report = runner.race_check(src, call, n * 4, n * 8, 1, kc["work_group_size"])
    assert ok, report[-3000:]


# ---- planted mistakes: the check must fire --------------------------------------------------

def test_missing_stage_barrier_is_reported():
    src, call, ib, ob, grid, block, _ = _plain("scfo8.256*4")
    bad = src.replace("__syncthreads();", ";", 1)
    ok, report = runner.race_check(bad, call, ib, ob, grid, block)
    assert not ok and "data race" in report


def test_missing_barrier_between_fused_passes_is_reported(monkeypatch):
    monkeypatch.setenv("BBFFT_FUSE", "on")
    src, call, ib, ob, grid, block = _fused(1, [16, 64], 9, 1)
    # the barrier between pass a (writes the plane) and pass b (reads it) in the plane loop
    loop = src.rindex("for (; f < planes;")
    cut = src.index("BBFFT_LDS_SYNC();", src.index("_a((const", loop))
    bad = src[:cut] + ";" + src[cut + len("BBFFT_LDS_SYNC();"):]
    ok, report = runner.race_check(bad, call, ib, ob, grid, block)
    assert not ok and "data race" in report


def test_missing_barrier_in_ragged_kernel_is_reported():
    jobs = ragged.pack([ragged.Job(M=2, N=64, K=3, direction=-1),
                        ragged.Job(M=1, N=128, K=2, direction=1)])
    src, call, ib, ob, grid, block, inplace, table = _ragged(jobs, grid=2)
    bad = src.replace("__syncthreads();", ";", 1)
    ok, report = runner.race_check(bad, call, ib, ob, grid, block, table=table)
    assert not ok and "data race" in report
