"""Persistent grids (work-groups per CU looping over the tiles) with and without the
software-pipelined tile loop (fft1d flag PREFETCH) on the GPU, through the C ABI: the result
matches the FP64 oracle and is bitwise equal to the one-tile-per-group kernel with the same
tiling (the arithmetic is the same; only the order of the memory accesses changes)."""
import numpy as np
import pytest

import bbfft_amd as bb
import oracle
from gpu_util import execute, have_gpu, io_shapes

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not have_gpu(), reason="needs a HIP GPU")]


def _input(cfg, seed):
    ishape, _ = io_shapes(cfg)
    n_in = oracle.extent(cfg.dim, ishape, cfg.istride)
    N = cfg.shape[1]
    rng = np.random.default_rng(seed)
    if int(cfg.type) == 1:
        return rng.uniform(-1, 1, n_in)
    return rng.uniform(-1, 1, n_in) + 1j * rng.uniform(-1, 1, n_in)


@pytest.mark.parametrize("desc", ["scfo8.256*3001", "dcfo8.256*2000", "scfo8.512*1500",
                                  "scbo5.60*2049", "dcfo3.105*1700", "srfo4.64*2500",
                                  "drfo2.27*3333", "scfi8.128*4096", "srfi1.100*5000"])
@pytest.mark.parametrize("groups", [1, 3])
def test_prefetch_persistent_grid(desc, groups):
    cfg = bb.parse_fft_descriptor(desc)
    base = bb.kernel_configuration(cfg)  # table / heuristic tiling, K tile of one column
    tuning = {"factorization": base["factorization"], "Nt": base["Nt"], "Mb": base["Mb"],
              "Kb": 1, "lds_pad": base["lds_pad"], "no_pad": base["lds_pad"] == 0, "flags": 0}
    x = _input(cfg, len(desc) * 7 + groups)
    inplace = desc[3] == "i"
    cache = bb.default_cache()
    ref_plan = bb.make_plan(cfg, cache=cache, tuning=tuning)
    y0 = execute(cfg, x, inplace=inplace, plan=ref_plan)
    for flags in (0, bb.PREFETCH):
        t = dict(tuning, flags=flags, groups_per_cu=groups)
        plan = bb.make_plan(cfg, cache=cache, tuning=t)
        (launch,) = plan.launches()
        kcols = (cfg.shape[2] + 1) // 2 if desc[1] == "r" and cfg.shape[1] % 2 else cfg.shape[2]
        tiles = -(-cfg.shape[0] // base["Mb"]) * kcols
        assert launch["grid"] < tiles  # every group walks several tiles
        assert launch["kernel"] == bb.kernel_configuration(cfg, tuning=t)["name"]
        y = execute(cfg, x, inplace=inplace, plan=plan)
        assert np.array_equal(y.view(np.uint8), y0.view(np.uint8))
    _, oshape = io_shapes(cfg)
    n_out = oracle.extent(cfg.dim, oshape, cfg.ostride)
    ref = oracle.fft(1, cfg.shape, cfg.istride, cfg.ostride, int(cfg.type), int(cfg.dir), x,
                     out_size=n_out)
    N = cfg.shape[1]
    mask = np.zeros(n_out, bool)
    M, K = cfg.shape[0], cfg.shape[2]
    H = N // 2 + 1 if int(cfg.type) == 1 else N
    m, n, k = np.meshgrid(np.arange(M), np.arange(H), np.arange(K), indexing="ij")
    mask[(m * cfg.ostride[0] + n * cfg.ostride[1] + k * cfg.ostride[2]).ravel()] = True
    assert oracle.rel_l2(y0[:n_out][mask], ref[mask]) <= oracle.tolerance(int(cfg.fp), N)


def test_table_persistent_grid_entry():
    """A table entry with groups_per_cu (fp64 N=512 M=8: one work-group per CU walking the
    tiles) launches that grid by default."""
    import torch
    cfg = bb.parse_fft_descriptor("dcfo8.512*32768")
    kc = bb.kernel_configuration(cfg)
    (launch,) = bb.make_plan(cfg, cache=bb.default_cache()).launches()
    assert launch["kernel"] == kc["name"]
    assert launch["grid"] == torch.cuda.get_device_properties(0).multi_processor_count
