"""Differentiation on the GPU: the adjoint kernels behind torch.autograd (bbfft_amd.autograd)
checked against FP64 finite differences (torch.autograd.gradcheck, and central differences on
the 3-D filter pipeline) and against torch.fft's gradients."""

import numpy as np
import pytest

from gpu_util import have_gpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not have_gpu(), reason="needs a HIP GPU")]

if have_gpu():
    import torch

    from bbfft_amd import autograd as ag


@pytest.mark.parametrize("shape,ndim", [((3, 16), 1), ((2, 15), 1), ((4, 6, 10), 2),
                                        ((2, 4, 6, 5), 3)])
def test_gradcheck_c2c(shape, ndim):
    x = torch.randn(*shape, dtype=torch.complex128, device="cuda", requires_grad=True)
    for d in (-1, 1):
        assert torch.autograd.gradcheck(lambda v: ag.fftn(v, ndim, d), (x,), eps=1e-6,
                                        atol=1e-8, rtol=1e-6)


@pytest.mark.parametrize("shape,ndim", [((3, 16), 1), ((2, 15), 1), ((4, 6, 10), 2),
                                        ((3, 5, 7), 2), ((2, 4, 6, 8), 3)])
def test_gradcheck_r2c_and_c2r(shape, ndim):
    x = torch.randn(*shape, dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda v: ag.rfftn(v, ndim), (x,), eps=1e-6, atol=1e-8,
                                    rtol=1e-6)
    n1 = shape[-1]
    hshape = list(shape[:-1]) + [n1 // 2 + 1]
    X = torch.randn(*hshape, dtype=torch.complex128, device="cuda", requires_grad=True)
    # c2r is differentiated as the map it computes, including the DC / Nyquist conventions
    assert torch.autograd.gradcheck(lambda v: ag.irfftn(v, n1, ndim), (X,), eps=1e-6,
                                    atol=1e-8, rtol=1e-6)


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_gradients_match_torch_fft(dtype):
    rdt = torch.float32 if dtype == "f32" else torch.float64
    cdt = torch.complex64 if dtype == "f32" else torch.complex128
    tol = 2e-5 if dtype == "f32" else 1e-12
    x = torch.randn(5, 12, 20, dtype=rdt, device="cuda", requires_grad=True)
    f = torch.randn(12, 11, dtype=cdt, device="cuda")
    f[:, 0] = f[:, 0].real.to(cdt)
    f[:, -1] = f[:, -1].real.to(cdt)
    f[1:, 0] = (f[1:, 0] + f[1:, 0].flip(0)) / 2   # Hermitian (real, symmetric) edge columns
    f[1:, -1] = (f[1:, -1] + f[1:, -1].flip(0)) / 2
    ours = (ag.irfftn(ag.rfftn(x, 2) * f, 20, 2) ** 2).sum()
    g_ours, = torch.autograd.grad(ours, x)
    xr = x.detach().clone().requires_grad_(True)
    ref = (torch.fft.irfftn(torch.fft.rfftn(xr, dim=(-2, -1)) * f, s=(12, 20), dim=(-2, -1),
                            norm="forward") ** 2).sum()
    g_ref, = torch.autograd.grad(ref, xr)
    assert abs(ours.item() / ref.item() - 1) < tol * 10
    assert (torch.linalg.vector_norm(g_ours - g_ref) / torch.linalg.vector_norm(g_ref)).item() < tol
    z = torch.randn(7, 64, dtype=cdt, device="cuda", requires_grad=True)
    g1, = torch.autograd.grad((ag.fftn(z, 1, -1).abs() ** 2).sum(), z)
    z2 = z.detach().clone().requires_grad_(True)
    g2, = torch.autograd.grad((torch.fft.fft(z2, dim=-1).abs() ** 2).sum(), z2)
    assert (torch.linalg.vector_norm(g1 - g2) / torch.linalg.vector_norm(g2)).item() < tol


def _filter(n=128, seed=7):
    """Seeded complex filter (n, n, n/2+1) whose k1 = 0 and k1 = n/2 planes are Hermitian in
    (k3, k2), so r2c -> filter -> c2r is the convolution-type map numpy/torch also compute."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    h = n // 2 + 1
    f = torch.complex(torch.randn(n, n, h, generator=g, device="cuda"),
                      torch.randn(n, n, h, generator=g, device="cuda"))
    for k1 in (0, h - 1):
        p = f[:, :, k1]
        pm = torch.roll(torch.flip(p, (0, 1)), shifts=(1, 1), dims=(0, 1))  # p[-k3, -k2]
        f[:, :, k1] = (p + pm.conj()) / 2
    return f


def _pipeline(x, f, n=128):
    X = ag.rfftn(x, 3)
    z = ag.irfftn(X * f, n, 3) / n ** 3
    return (z * z).sum(), z


def test_3d_filter_pipeline_gradient():
    """32 x 128^3 fp32 real input: r2c -> spectral filter -> c2r / N^3, loss = sum of squares.
    The adjoint-kernel gradient is checked against FP64 central differences on 64 random
    entries (the loss is quadratic, so central differences are exact up to rounding) and
    against torch.fft's gradient; the backward/forward time ratio is reported."""
    n, B = 128, 32
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    x = torch.randn(B, n, n, n, generator=g, device="cuda", dtype=torch.float32,
                    requires_grad=True)
    f = _filter(n)
    loss, _ = _pipeline(x, f)
    grad, = torch.autograd.grad(loss, x)
    # reference gradient: torch.fft autograd in fp64 on the same input
    x64 = x.detach().double().requires_grad_(True)
    f64 = f.to(torch.complex128)
    z64 = torch.fft.irfftn(torch.fft.rfftn(x64, dim=(1, 2, 3)) * f64, s=(n, n, n),
                           dim=(1, 2, 3), norm="forward") / n ** 3
    g64, = torch.autograd.grad((z64 * z64).sum(), x64)
    rel = (torch.linalg.vector_norm(grad.double() - g64) / torch.linalg.vector_norm(g64)).item()
    assert rel < 1e-5, rel
    # FP64 central differences on 64 entries (loss evaluated in fp64 with torch.fft)
    rng = np.random.default_rng(5)
    idx = [tuple(int(v) for v in rng.integers(0, s)) for s in [(B, n, n, n)] * 64]
    h = 1e-3

    def loss64(v):
        zz = torch.fft.irfftn(torch.fft.rfftn(v, dim=(1, 2, 3)) * f64, s=(n, n, n),
                              dim=(1, 2, 3), norm="forward") / n ** 3
        return (zz * zz).sum().item()

    xd = x.detach().double()
    fd, an = [], []
    with torch.no_grad():
        for i in idx:
            old = xd[i].item()
            xd[i] = old + h
            lp = loss64(xd)
            xd[i] = old - h
            lm = loss64(xd)
            xd[i] = old
            fd.append((lp - lm) / (2 * h))
            an.append(grad[i].item())
    fd, an = np.array(fd), np.array(an)
    assert np.linalg.norm(an - fd) / np.linalg.norm(fd) < 1e-4


def test_3d_backward_forward_time_ratio():
    """The hand-written adjoint keeps the backward pass within the 1.2x-of-forward target on
    the 32 x 128^3 filter pipeline (median of 20 event-timed repetitions each)."""
    n, B = 128, 32
    x = torch.randn(B, n, n, n, device="cuda", dtype=torch.float32, requires_grad=True)
    f = _filter(n)
    for _ in range(3):
        loss, _ = _pipeline(x, f)
        torch.autograd.grad(loss, x)
    torch.cuda.synchronize()
    tf, tb = [], []
    for _ in range(20):
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record()
        loss, _ = _pipeline(x, f)
        e1.record()
        torch.autograd.grad(loss, x)
        e2.record()
        torch.cuda.synchronize()
        tf.append(e0.elapsed_time(e1))
        tb.append(e1.elapsed_time(e2))
    mf, mb = float(np.median(tf)), float(np.median(tb))
    ratio = mb / mf
    print(f"3-D filter pipeline: forward {mf:.3f} ms, backward {mb:.3f} ms, ratio {ratio:.3f}")
    assert ratio <= 1.2


def test_filtered_irfftn_gradcheck():
    """irfftn with the spectral product and scale fused into its first pass (load callback)
    and conj(filter) * scale into its adjoint's last pass (store callback): FP64 gradcheck."""
    shape, n1 = (2, 4, 6, 5), 8
    X = torch.randn(2, 4, 6, n1 // 2 + 1, dtype=torch.complex128, device="cuda",
                    requires_grad=True)
    F = torch.randn(4, 6, n1 // 2 + 1, dtype=torch.complex128, device="cuda")
    assert torch.autograd.gradcheck(lambda v: ag.irfftn(v, n1, 3, filter=F, scale=0.25),
                                    (X,), eps=1e-6, atol=1e-8, rtol=1e-6)
    Y = torch.randn(3, 10, 7, dtype=torch.complex128, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda v: ag.irfftn(v, 12, 2, scale=3.0), (Y,), eps=1e-6,
                                    atol=1e-8, rtol=1e-6)


def _pipeline_fused(x, f, n=128):
    z = ag.irfftn(ag.rfftn(x, 3), n, 3, filter=f, scale=1.0 / n ** 3)
    return (z * z).sum(), z


def test_3d_filter_pipeline_fused_matches_composition():
    """32 x 128^3 fp32: the fused spectral product gives the composition's values and
    gradients (the product and scale are rounded in another order)."""
    n, B = 128, 32
    g = torch.Generator(device="cuda")
    g.manual_seed(13)
    x = torch.randn(B, n, n, n, generator=g, device="cuda", dtype=torch.float32,
                    requires_grad=True)
    f = _filter(n)
    loss_a, z_a = _pipeline(x, f)
    loss_b, z_b = _pipeline_fused(x, f)
    rel = (torch.linalg.vector_norm(z_a - z_b) / torch.linalg.vector_norm(z_a)).item()
    assert rel < 1e-6, rel
    ga, = torch.autograd.grad(loss_a, x)
    gb, = torch.autograd.grad(loss_b, x)
    rel = (torch.linalg.vector_norm(ga - gb) / torch.linalg.vector_norm(ga)).item()
    assert rel < 1e-6, rel


def test_3d_fused_backward_forward_time_ratio():
    """The fused pipeline also keeps backward within 1.2x of forward (median of 20)."""
    n, B = 128, 32
    x = torch.randn(B, n, n, n, device="cuda", dtype=torch.float32, requires_grad=True)
    f = _filter(n)
    for _ in range(3):
        loss, _ = _pipeline_fused(x, f)
        torch.autograd.grad(loss, x)
    torch.cuda.synchronize()
    tf, tb = [], []
    for _ in range(20):
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record()
        loss, _ = _pipeline_fused(x, f)
        e1.record()
        torch.autograd.grad(loss, x)
        e2.record()
        torch.cuda.synchronize()
        tf.append(e0.elapsed_time(e1))
        tb.append(e1.elapsed_time(e2))
    mf, mb = float(np.median(tf)), float(np.median(tb))
    print(f"fused 3-D filter pipeline: forward {mf:.3f} ms, backward {mb:.3f} ms, "
          f"ratio {mb / mf:.3f}")
    assert mb / mf <= 1.2
