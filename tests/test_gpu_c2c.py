"""GPU parity of complex transforms (reference: test/c2c.cpp) through the C ABI."""
import zlib

import numpy as np
import pytest

import bbfft_amd as bb
import gpu_cases
import oracle
from gpu_util import approx_equal, execute, grid, have_gpu, offsets, ref_tol

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not have_gpu(), reason="needs a HIP GPU")]

TAU = 2.0 * np.pi


@pytest.mark.parametrize("desc", gpu_cases.c2c_known_answer())
def test_c2c_forward_plane_wave(desc):
    """x = s(k) e^{+2 pi i b n / N} / N with b = (m + k) mod N -> X = s(k) delta(n - b)
    (test/c2c.cpp:17-53), in place."""
    cfg = bb.parse_fft_descriptor(desc)
    M, N, K = cfg.shape[:3]
    m, n, k = grid([M, N, K])
    b = (m + k) % N
    scale = 1.0 + k / K
    x = np.zeros(offsets([m, n, k], cfg.istride[:3]).max() + 1, dtype=np.complex128)
    x[offsets([m, n, k], cfg.istride[:3])] = scale * np.exp(1j * TAU * ((b * n) % N) / N) / N
    y = execute(cfg, x, inplace=True)
    out = y[offsets([m, n, k], cfg.ostride[:3])]
    ref = scale * ((n - b) % N == 0)
    eps = ref_tol(cfg.fp, N)
    assert approx_equal(out.real, ref, eps).all()
    assert approx_equal(out.imag, 0.0, eps).all()


@pytest.mark.parametrize("desc", gpu_cases.c2c_nonpacked())
def test_c2c_nonpacked_forward(desc):
    """Strides {1, M+1, (M+1)(N+1)} (test/c2c.cpp:66-83)."""
    cfg = bb.parse_fft_descriptor(desc)
    M, N, K = cfg.shape[:3]
    m, n, k = grid([M, N, K])
    b = (m + k) % N
    scale = 1.0 + k / K
    off = offsets([m, n, k], cfg.istride[:3])
    x = np.zeros(off.max() + 1, dtype=np.complex128)
    x[off] = scale * np.exp(1j * TAU * ((b * n) % N) / N) / N
    y = execute(cfg, x, inplace=True)
    out = y[offsets([m, n, k], cfg.ostride[:3])]
    ref = scale * ((n - b) % N == 0)
    eps = ref_tol(cfg.fp, N)
    assert approx_equal(out.real, ref, eps).all()
    assert approx_equal(out.imag, 0.0, eps).all()


@pytest.mark.parametrize("desc", [d for d in gpu_cases.c2c_identity() if d[2] == "f"])
def test_c2c_identity(desc):
    """backward(forward(x)) / N == x (test/c2c.cpp:85-127), seeded instead of random_device."""
    cfg = bb.parse_fft_descriptor(desc)
    icfg = bb.parse_fft_descriptor(desc.replace("cfi", "cbi"))
    M, N, K = cfg.shape[:3]
    x = oracle.uniform_complex(11, M * N * K)
    X = execute(cfg, x, inplace=True)
    xr = execute(icfg, X, inplace=True) / N
    eps = ref_tol(cfg.fp, N)
    assert approx_equal(xr.real, x.real, eps).all()
    assert approx_equal(xr.imag, x.imag, eps).all()


@pytest.mark.parametrize("desc", gpu_cases.c1_full_k())
def test_c1_shape_full_k_vs_oracle(desc):
    """M=1 x N=64 x K=16384 at full K, seeded uniform[-1,1) inputs, against the FP64 oracle."""
    cfg = bb.parse_fft_descriptor(desc)
    M, N, K = cfg.shape[:3]
    x = oracle.uniform_complex(1, M * N * K)
    y = execute(cfg, x)
    ref = oracle.fft(1, cfg.shape, cfg.istride, cfg.ostride, 0, int(cfg.dir), x)
    assert oracle.rel_l2(y, ref) <= oracle.tolerance(int(cfg.fp), N)


@pytest.mark.parametrize("desc", gpu_cases.c2c_random())
def test_c2c_random_vs_oracle(desc):
    """Seeded uniform[-1,1) inputs vs the FP64 DFT oracle: rel-L2 <= 8 eps log2 N."""
    cfg = bb.parse_fft_descriptor(desc)
    M, N, K = cfg.shape[:3]
    x = oracle.uniform_complex(zlib.crc32(desc.encode()) & 0xFFFF, M * N * K)
    y = execute(cfg, x)
    ref = oracle.fft(1, cfg.shape, cfg.istride, cfg.ostride, 0, int(cfg.dir), x)
    assert oracle.rel_l2(y, ref) <= oracle.tolerance(int(cfg.fp), N)


@pytest.mark.parametrize("desc", gpu_cases.c2c_full_size())
def test_c2c_full_size(desc):
    """Benchmark shape M=8 x N x K=32768: 256 sampled columns vs the oracle, the full tensor
    through backward(forward(x)) / N, and Parseval over the whole tensor."""
    import torch
    cfg = bb.parse_fft_descriptor(desc)
    M, N, K = cfg.shape[:3]
    fp = int(cfg.fp)
    cdt = torch.complex64 if fp == 4 else torch.complex128
    g = torch.Generator(device="cuda")
    g.manual_seed(1234 + N)
    rdt = torch.float32 if fp == 4 else torch.float64
    x = torch.view_as_complex(torch.rand((K, N, M, 2), generator=g, device="cuda", dtype=rdt)
                              * 2 - 1).contiguous()
    y = torch.empty_like(x)
    plan = bb.make_plan(cfg, cache=bb.default_cache())
    plan.execute(x, y)
    inv = bb.parse_fft_descriptor(desc.replace("cfo", "cbo") if "cfo" in desc
                                  else desc.replace("cbo", "cfo"))
    iplan = bb.make_plan(inv, cache=bb.default_cache())
    z = torch.empty_like(x)
    iplan.execute(y, z)
    torch.cuda.synchronize()
    tol = oracle.tolerance(fp, N)
    # identity over the full tensor
    err = (torch.linalg.vector_norm(z / N - x) / torch.linalg.vector_norm(x)).item()
    assert err <= 2 * tol
    # Parseval: ||X||^2 = N ||x||^2
    px = torch.linalg.vector_norm(x.to(torch.complex128)).item() ** 2
    py = torch.linalg.vector_norm(y.to(torch.complex128)).item() ** 2
    assert abs(py / (N * px) - 1.0) <= 2 * tol
    # sampled columns vs oracle
    rng = np.random.default_rng(N)
    ks = torch.tensor(np.sort(rng.choice(K, 32, replace=False)), device="cuda")
    xs = x[ks].cpu().numpy().astype(np.complex128)  # (32, N, M)
    ys = y[ks].cpu().numpy().astype(np.complex128)
    ref = oracle.fft(1, [M, N, 32], [1, M, M * N], [1, M, M * N], 0, int(cfg.dir),
                     xs.ravel())
    assert oracle.rel_l2(ys.ravel(), ref) <= tol


def test_plan_reuse_and_events():
    """execute() waits for `wait` events and records `signal` on the plan's stream."""
    import torch
    cfg = bb.parse_fft_descriptor("scfo8.64*100")
    plan = bb.make_plan(cfg, cache=bb.default_cache())
    x = torch.view_as_complex(torch.randn(100, 64, 8, 2, device="cuda")).contiguous()
    y1, y2 = torch.empty_like(x), torch.empty_like(x)
    e_in = torch.cuda.Event()
    e_out = torch.cuda.Event()
    e_out.record()  # creates the underlying hipEvent_t
    e_in.record()
    plan.execute(x, y1, signal=e_out.cuda_event, wait=[e_in.cuda_event])
    e_out.synchronize()
    plan.execute(x, y2)
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)
    ref = torch.fft.fft(x, dim=1)
    assert (torch.linalg.vector_norm(y1 - ref) / torch.linalg.vector_norm(ref)).item() < 1e-5


def test_launch_introspection():
    cfg = bb.parse_fft_descriptor("scfo8.256*32768")
    kc = bb.kernel_configuration(cfg)
    plan = bb.make_plan(cfg, cache=bb.default_cache())
    (l,) = plan.launches()
    assert l["kernel"] == kc["name"] and l["kernel"].startswith("bbfft_m1_M8_Mb8_N256")
    assert l["block"] == kc["work_group_size"]
    assert l["grid"] == 32768 // kc["Kb"]


def test_execute_rejects_small_device_buffers():
    """Tensor arguments smaller than the configured tensor are refused before any launch (the
    kernel would address them out of bounds); exact-size tensors and views into larger
    storage run."""
    import torch
    cfg = bb.parse_fft_descriptor("scfo8.256*64")
    plan = bb.make_plan(cfg, cache=bb.default_cache())
    n = 8 * 256 * 64
    x = torch.zeros(n, dtype=torch.complex64, device="cuda")
    y = torch.zeros(n, dtype=torch.complex64, device="cuda")
    small = torch.zeros(n - 1, dtype=torch.complex64, device="cuda")
    for args in ((small, y), (x, small), (small,)):
        with pytest.raises(ValueError, match="holds"):
            plan.execute(*args)
    big = torch.zeros(n + 16, dtype=torch.complex64, device="cuda")
    plan.execute(x, big[16:])
    plan.execute(big[:n])
    torch.cuda.synchronize()
