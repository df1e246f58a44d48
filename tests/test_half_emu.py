"""Half-precision input kernels on the CPU (fiber emulator, clang host build): the plain
variant (convert at load) and the MFMA variant (radix-16 stage as a split DFT-matrix GEMM,
matrix instructions emulated with the documented operand layout) against the FP64 oracle on
the dequantised input; the split DFT-matrix terms checked against the exact matrix."""
import re

import numpy as np
import pytest

import bbfft_amd as bb
import half_util
import oracle
from emu import runner


def _cfg(M, N, K, d=-1):
    cfg = bb.configuration(1, [M, N, K], bb.precision.f32, bb.direction(d),
                           bb.transform_type.c2c)
    cfg.set_strides_default(False)
    return cfg


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("mfma", [False, True])
@pytest.mark.parametrize("M,N,K,d", [(1, 32, 3, -1), (4, 512, 2, -1), (3, 48, 5, 1),
                                     (2, 256, 3, 1), (1, 80, 4, -1), (8, 64, 3, -1)])
def test_half_input_vs_oracle(storage, mfma, M, N, K, d):
    rng = np.random.default_rng(N * 31 + M)
    z = rng.uniform(-1, 1, M * N * K) + 1j * rng.uniform(-1, 1, M * N * K)
    bits = half_util.pack(z, storage)
    x = half_util.unpack(bits, storage)
    cfg = _cfg(M, N, K, d)
    flags = bb.half_input_flags(storage, mfma)
    y = runner.run_half_input(bb, cfg, bits, flags, grid_cap=2)
    ref = oracle.fft(1, cfg.shape, cfg.istride, cfg.ostride, 0, d, x)
    assert oracle.rel_l2(y, ref) <= oracle.tolerance(4, N)


def test_heavy_tail_mfma():
    rng = np.random.default_rng(3)
    M, N, K = 4, 512, 2
    for storage in ("f16", "bf16"):
        bits = half_util.pack(half_util.heavy_tail(rng, M * N * K), storage)
        x = half_util.unpack(bits, storage)
        cfg = _cfg(M, N, K)
        y = runner.run_half_input(bb, cfg, bits, bb.half_input_flags(storage, True))
        ref = oracle.fft(1, cfg.shape, cfg.istride, cfg.ostride, 0, -1, x)
        assert oracle.rel_l2(y, ref) <= 1e-6, storage


@pytest.mark.parametrize("storage,terms,bits", [("f16", 2, 22), ("bf16", 3, 24)])
@pytest.mark.parametrize("d", [-1, 1])
def test_split_matrix_terms(storage, terms, bits, d):
    """The emitted table holds `terms` half-precision matrices whose sum equals the real 32 x 32
    form of the 16-point DFT matrix to within 2^-bits (relative to 1)."""
    cfg = _cfg(4, 512, 1, d)
    src = bb.generate_source(cfg, tuning={"flags": bb.half_input_flags(storage, True)})
    m = re.search(r"_mA\[(\d+)\] = \{(.*?)\};", src, re.S)
    vals = np.array([int(v) for v in re.findall(r"(\d+)u", m.group(2))], dtype=np.uint32)
    assert int(m.group(1)) == terms * 1024 == vals.size
    if storage == "f16":
        t = vals.astype(np.uint16).view(np.float16).astype(np.float64)
    else:
        t = (vals << 16).view(np.float32).astype(np.float64)
    A = t.reshape(terms, 32, 32).sum(axis=0)
    k1 = np.arange(16)[:, None]
    n1 = np.arange(16)[None, :]
    W = np.exp(d * 2j * np.pi * k1 * n1 / 16)
    ref = np.zeros((32, 32))
    ref[0::2, 0::2] = W.real
    ref[0::2, 1::2] = -W.imag
    ref[1::2, 0::2] = W.imag
    ref[1::2, 1::2] = W.real
    assert np.max(np.abs(A - ref)) <= 2.0 ** -bits
    # the leading term alone is the matrix rounded to the half format
    lead = t.reshape(terms, 32, 32)[0]
    if storage == "f16":
        assert np.array_equal(lead, ref.astype(np.float16).astype(np.float64))


def test_half_input_validation():
    for desc, flags in [("dcfo4.512*2", bb.INPUT_F16), ("srfo4.512*2", bb.INPUT_F16),
                        ("scfo4.24*2", bb.INPUT_F16 | bb.MFMA),
                        ("scfo4.16*2", bb.INPUT_F16 | bb.MFMA),
                        ("scfo4.512*2", bb.MFMA),
                        ("scfo4.512*2", bb.INPUT_F16 | bb.INPUT_BF16)]:
        with pytest.raises(bb.BadConfiguration):
            bb.kernel_configuration(bb.parse_fft_descriptor(desc), tuning={"flags": flags})
    kc = bb.kernel_configuration(bb.parse_fft_descriptor("scfo4.512*262144"),
                                 tuning={"flags": bb.half_input_flags("f16", True)})
    assert kc["factorization"][0] == 16 and kc["work_group_size"] <= 1024
