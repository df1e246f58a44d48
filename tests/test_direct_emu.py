"""Direct DFT stages (radices above the in-register limit, e.g. the prime factors of N = 67,
134, 254, 262, 509) compiled for the host and run by the fiber emulator against the FP64
oracle. The reference plans these lengths with its basic_esum fallback
(src/base/mixed_radix_fft.cpp:46-53; factor2_slm_fft.cpp:34-40 stops factoring at a prime)."""
import numpy as np
import pytest

import bbfft_amd as bb
import oracle
from emu import runner

LENGTHS = [67, 134, 254, 262, 393, 469, 509]


def _cfg(M, N, K, fp=4, d=-1, t=0, inplace=False, dims=None):
    dims = dims or [N]
    cfg = bb.configuration(len(dims), [M] + dims + [K], bb.precision(fp), bb.direction(d),
                           bb.transform_type(t))
    cfg.set_strides_default(inplace)
    return cfg


@pytest.mark.parametrize("N", LENGTHS)
def test_lengths_plan_a_direct_stage(N):
    kc = bb.kernel_configuration(bb.parse_fft_descriptor(f"scfo8.{N}*32768"))
    big = [f for f in kc["factorization"] if f > 64]
    assert len(big) == 1 and big[0] % 2 == 1
    assert int(np.prod(kc["factorization"])) == N
    assert kc["Nt"] % (N // big[0]) == 0 and kc["work_group_size"] <= 1024


@pytest.mark.parametrize("fp", [4, 8])
@pytest.mark.parametrize("N", LENGTHS)
def test_c2c(N, fp):
    for M, K, d, cap in [(3, 5, -1, 2), (2, 3, 1, None)]:
        cfg = _cfg(M, N, K, fp, d)
        x = oracle.uniform_complex(N + 1, M * N * K)
        y = runner.run(bb, cfg, x, M * N * K, grid_cap=cap)
        ref = oracle.fft(1, cfg.shape, cfg.istride, cfg.ostride, 0, d, x)
        assert oracle.rel_l2(y, ref) <= oracle.tolerance(fp, N)


@pytest.mark.parametrize("fac", [[2, 67], [67, 2], [2, 131], [7, 67], [67, 7]])
def test_direct_stage_at_any_position(fac):
    """The direct stage after other stages applies the Stockham twiddles in its input phase."""
    N = int(np.prod(fac))
    for fp in (4, 8):
        cfg = _cfg(2, N, 3, fp, -1)
        R = max(fac)
        tuning = {"factorization": fac, "Nt": (N // R) * 3}
        x = oracle.uniform_complex(N + 7, 2 * N * 3)
        y = runner.run(bb, cfg, x, 2 * N * 3, tuning=tuning)
        ref = oracle.fft(1, cfg.shape, cfg.istride, cfg.ostride, 0, -1, x)
        assert oracle.rel_l2(y, ref) <= oracle.tolerance(fp, N)


@pytest.mark.parametrize("N", [67, 134, 254, 262, 509])
@pytest.mark.parametrize("fp", [4, 8])
def test_r2c(N, fp):
    M, K = 2, 5
    cfg = _cfg(M, N, K, fp, -1, 1)
    n_in = oracle.extent(1, cfg.shape, cfg.istride)
    n_out = oracle.extent(1, [M, N // 2 + 1, K], cfg.ostride)
    x = oracle.uniform_real(N + 2, n_in)
    ref = oracle.fft(1, cfg.shape, cfg.istride, cfg.ostride, 1, -1, x, out_size=n_out)
    y = runner.run(bb, cfg, x, n_out, grid_cap=2)
    assert oracle.rel_l2(y, ref) <= oracle.tolerance(fp, N)


@pytest.mark.parametrize("N", [67, 134, 254, 262, 509])
@pytest.mark.parametrize("fp", [4, 8])
def test_c2r(N, fp):
    """Non-zero DC and Nyquist imaginary parts (reference rules)."""
    M, K = 2, 5
    cfg = _cfg(M, N, K, fp, 1, 2)
    n_in = oracle.extent(1, [M, N // 2 + 1, K], cfg.istride)
    n_out = oracle.extent(1, cfg.shape, cfg.ostride)
    X = oracle.uniform_complex(N + 3, n_in)
    ref = oracle.fft(1, cfg.shape, cfg.istride, cfg.ostride, 2, 1, X, out_size=n_out)
    y = runner.run(bb, cfg, X, n_out, grid_cap=2)
    m, n, k = np.meshgrid(range(M), range(N), range(K), indexing="ij")
    sel = (m + n * cfg.ostride[1] + k * cfg.ostride[2]).ravel()
    assert oracle.rel_l2(y[sel], ref[sel]) <= oracle.tolerance(fp, N)


def test_ragged_rejects_direct_stage_lengths():
    from bbfft_amd import ragged
    jobs = ragged.pack([ragged.Job(M=1, N=64, K=4), ragged.Job(M=1, N=134, K=4)])
    with pytest.raises(bb.BadConfiguration):
        ragged.kernel_configuration(jobs)


def test_two_direct_factors_rejected():
    with pytest.raises(bb.BadConfiguration):
        bb.kernel_configuration(bb.parse_fft_descriptor("scfo1.4757*2"))  # 67 * 71
