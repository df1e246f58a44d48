"""GPU parity at lengths whose prime factor exceeds the in-register radix limit: the direct
DFT stage (the reference's basic_esum fallback, src/base/mixed_radix_fft.cpp:46-53), through
the C ABI against the FP64 oracle."""
import zlib

import numpy as np
import pytest

import bbfft_amd as bb
import gpu_cases
import oracle
from gpu_util import execute, have_gpu, io_shapes

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not have_gpu(), reason="needs a HIP GPU")]


@pytest.mark.parametrize("desc", gpu_cases.direct_stage())
def test_direct_stage_vs_oracle(desc):
    """Seeded inputs (fully random half spectra for c2r) against the FP64 oracle."""
    cfg = bb.parse_fft_descriptor(desc)
    t = int(cfg.type)
    ishape, _ = io_shapes(cfg)
    n_in = oracle.extent(cfg.dim, ishape, cfg.istride)
    seed = zlib.crc32(desc.encode()) & 0xFFFF
    x = oracle.uniform_real(seed, n_in) if t == 1 else oracle.uniform_complex(seed, n_in)
    y = execute(cfg, x)
    ref = oracle.fft(cfg.dim, cfg.shape, cfg.istride, cfg.ostride, t, int(cfg.dir), x,
                     out_size=y.size)
    n_total = int(np.prod(cfg.shape[1:cfg.dim + 1]))
    assert oracle.rel_l2(y, ref) <= oracle.tolerance(int(cfg.fp), n_total)


@pytest.mark.parametrize("desc", gpu_cases.direct_full_size())
def test_direct_stage_full_size_round_trip_and_sampled_columns(desc):
    """M=8 x N=509 x K=32768: backward(forward(x)) = N x, and 64 sampled columns against the
    oracle."""
    cfg = bb.parse_fft_descriptor(desc)
    M, N, K = cfg.shape[:3]
    inverse = bb.parse_fft_descriptor(desc.replace("cfo", "cxo").replace("cbo", "cfo")
                                      .replace("cxo", "cbo"))
    x = oracle.uniform_complex(509, M * N * K)
    y = execute(cfg, x)
    z = execute(inverse, y)
    fp = int(cfg.fp)
    assert oracle.rel_l2(z / N, x) <= 2 * oracle.tolerance(fp, N)
    rng = np.random.default_rng(7)
    ks = rng.choice(K, 8, replace=False)
    xs = x.reshape(K, N, M)[ks]
    ref = oracle.fft(1, [M, N, 8], [1, M, M * N], [1, M, M * N], 0, int(cfg.dir), xs.ravel())
    got = y.reshape(K, N, M)[ks].ravel()
    assert oracle.rel_l2(got, ref) <= oracle.tolerance(fp, N)
