"""Half-precision input on the GPU: plain (convert at load) and MFMA (radix-16 stage as a split
DFT-matrix GEMM on the matrix cores) kernels against the FP64 oracle on the dequantised input;
the full 4 x 512 x 262144 workload (f16 / bf16 storage, uniform and heavy-tail inputs) against
an FP64 FFT of the same dequantised values, with relative-L2 <= 1e-6 and the errors printed."""
import numpy as np
import pytest

import bbfft_amd as bb
import half_util
import oracle
from gpu_util import have_gpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not have_gpu(), reason="needs a HIP GPU")]

if have_gpu():
    import torch


def _cfg(M, N, K, d=-1):
    cfg = bb.configuration(1, [M, N, K], bb.precision.f32, bb.direction(d),
                           bb.transform_type.c2c)
    cfg.set_strides_default(False)
    return cfg


def _run(cfg, bits, storage, mfma):
    xd = torch.from_numpy(bits.view(np.int32)).to("cuda")
    M, N, K = cfg.shape[:3]
    yd = torch.empty(M * N * K, dtype=torch.complex64, device="cuda")
    plan = bb.make_half_input_plan(cfg, storage, mfma, cache=bb.default_cache())
    plan.execute(xd, yd)
    torch.cuda.synchronize()
    return yd.cpu().numpy()


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("mfma", [False, True])
@pytest.mark.parametrize("M,N,K,d", [(1, 32, 3, -1), (4, 512, 40, -1), (3, 48, 25, 1),
                                     (2, 256, 17, 1), (1, 80, 64, -1), (8, 64, 9, -1),
                                     (5, 128, 33, 1), (1, 384, 12, -1), (4, 96, 70, 1)])
def test_half_input_vs_oracle(storage, mfma, M, N, K, d):
    rng = np.random.default_rng(N * 31 + M)
    z = rng.uniform(-1, 1, M * N * K) + 1j * rng.uniform(-1, 1, M * N * K)
    bits = half_util.pack(z, storage)
    x = half_util.unpack(bits, storage)
    cfg = _cfg(M, N, K, d)
    y = _run(cfg, bits, storage, mfma)
    ref = oracle.fft(1, cfg.shape, cfg.istride, cfg.ostride, 0, d, x)
    assert oracle.rel_l2(y, ref) <= oracle.tolerance(4, N)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_tuned_plan_with_half_input_bits(storage):
    """make_plan with a tuning whose flags carry the half-input (and MFMA) bits, as the
    autotuner builds its candidates: the plan takes a buffer of 4 bytes per input element and
    computes what the half-input plan with the same tiling computes."""
    M, N, K = 4, 512, 96
    cfg = _cfg(M, N, K)
    kc = bb.half_input_kernel_configuration(cfg, storage, True)
    flags = bb.half_input_flags(storage, True)
    tuning = {"factorization": kc["factorization"], "Nt": kc["Nt"], "Mb": kc["Mb"],
              "Kb": kc["Kb"], "lds_pad": kc["lds_pad"], "no_pad": kc["lds_pad"] == 0,
              "flags": flags, "waves_per_eu": 0, "groups_per_cu": 0}
    rng = np.random.default_rng(5)
    z = rng.uniform(-1, 1, M * N * K) + 1j * rng.uniform(-1, 1, M * N * K)
    bits = half_util.pack(z, storage)
    xd = torch.from_numpy(bits.view(np.int32)).to("cuda")  # exactly M N K words
    yd = torch.empty(M * N * K, dtype=torch.complex64, device="cuda")
    bb.make_plan(cfg, cache=bb.default_cache(), tuning=tuning).execute(xd, yd)
    torch.cuda.synchronize()
    ref = oracle.fft(1, cfg.shape, cfg.istride, cfg.ostride, 0, -1,
                     half_util.unpack(bits, storage))
    assert oracle.rel_l2(yd.cpu().numpy(), ref) <= oracle.tolerance(4, N)


def test_half_input_launch_and_alignment():
    cfg = _cfg(4, 512, 64)
    plan = bb.make_half_input_plan(cfg, "f16", True, cache=bb.default_cache())
    (launch,) = plan.launches()
    kc = bb.half_input_kernel_configuration(cfg, "f16", True)  # tuning table applied
    assert launch["kernel"] == kc["name"] and launch["block"] % 64 == 0
    x = torch.zeros(4 * 512 * 64 + 1, dtype=torch.int32, device="cuda")
    y = torch.zeros(4 * 512 * 64, dtype=torch.complex64, device="cuda")
    with pytest.raises(bb.BbfftError):
        plan.execute(x.data_ptr() + 2, y.data_ptr())


def _full_case(storage, dist, M=4, N=512, K=262144, seed=17):
    """Inputs drawn on the GPU (uniform[-1,1) or heavy tail), quantised to the storage format;
    returns (packed int32 device tensor, complex128 dequantised device tensor)."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    n = M * N * K
    if dist == "uniform":
        re = torch.rand(n, generator=g, device="cuda") * 2 - 1
        im = torch.rand(n, generator=g, device="cuda") * 2 - 1
    else:
        mag = 10.0 ** (torch.rand(n, generator=g, device="cuda", dtype=torch.float64) * 6 - 3)
        ph = torch.rand(n, generator=g, device="cuda", dtype=torch.float64) * (2 * np.pi)
        re = (mag * torch.cos(ph)).float()
        im = (mag * torch.sin(ph)).float()
        del mag, ph
    hdt = torch.float16 if storage == "f16" else torch.bfloat16
    pair = torch.stack([re.to(hdt), im.to(hdt)], dim=1).contiguous()
    del re, im
    x64 = torch.complex(pair[:, 0].double(), pair[:, 1].double())
    packed = pair.view(torch.int32).reshape(n)
    return packed, x64


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("dist", ["uniform", "heavy_tail"])
def test_full_size_mfma_error(storage, dist):
    """M=4 x N=512 x K=262144: MFMA-split result vs FP64 FFT (torch.fft on complex128) of the
    same dequantised input; relative-L2 <= 1e-6. 64 columns are also checked against the
    oracle."""
    M, N, K = 4, 512, 262144
    torch.cuda.empty_cache()
    packed, x64 = _full_case(storage, dist, M, N, K)
    cfg = _cfg(M, N, K)
    y = torch.empty(M * N * K, dtype=torch.complex64, device="cuda")
    plan = bb.make_half_input_plan(cfg, storage, True, cache=bb.default_cache())
    plan.execute(packed, y)
    torch.cuda.synchronize()
    # column-major M x N x K == row-major (K, N, M): transform along dim 1
    ref = torch.fft.fft(x64.view(K, N, M), dim=1).reshape(-1)
    diff = y.to(torch.complex128) - ref
    rel = (torch.linalg.vector_norm(diff) / torch.linalg.vector_norm(ref)).item()
    mx = diff.abs().max().item()
    print(f"{storage} {dist}: rel-L2 {rel:.3e}, max-abs {mx:.3e} "
          f"(max |X| {ref.abs().max().item():.3e})", flush=True)
    assert rel <= 1e-6
    cols = torch.arange(64, device="cuda") * (K // 64)
    xs = x64.view(K, N, M)[cols].cpu().numpy()  # (64, N, M)
    ys = y.view(K, N, M)[cols].cpu().numpy()
    xcm = np.ascontiguousarray(xs).reshape(-1)
    refo = oracle.fft(1, [M, N, 64], [1, M, M * N], [1, M, M * N], 0, -1, xcm)
    assert oracle.rel_l2(np.ascontiguousarray(ys).reshape(-1), refo) <= 1e-6
    del packed, x64, y, ref, diff
    torch.cuda.empty_cache()
