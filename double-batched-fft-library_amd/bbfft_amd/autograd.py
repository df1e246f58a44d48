"""Differentiable transforms for PyTorch over the gfx950 kernels (thin binding over the C ABI).

    X = autograd.rfftn(x, ndim=3)          # real (..., N3, N2, N1) -> complex (..., N3, N2, N1/2+1)
    x = autograd.irfftn(X, n1=128, ndim=3) # complex half spectrum -> real, unnormalised
    Y = autograd.fftn(x, ndim=1, direction=-1)

Tensors are torch tensors in row-major order whose trailing `ndim` axes are the FFT modes (the
last axis is bbfft's N_1, the fastest mode) and whose leading axes are flattened into bbfft's K
batch mode (M = 1). Transforms are unnormalised, like the reference. The backward pass runs the
adjoint kernels (bbfft_plan_create_adjoint): the opposite-sign transform for c2c, the
Hermitian-fold weighting + c2r for r2c, and r2c + doubling of the non-self-conjugate bins for
c2r. The c2r forward keeps the reference's conventions (imag of the DC bin ignored; for even N_1
the Nyquist imaginary part contributes -imag(X[N/2]) to every sample), and its backward is the
exact adjoint of that map.
"""
import math

import torch

from . import configuration, default_cache, direction, make_adjoint_plan, make_plan, precision
from . import transform_type

_plans = {}


def _cfg(kind, fft_shape, batch, fp, dirn):
    """bbfft configuration for contiguous row-major tensors; fft_shape is (N1, ..., Nd)."""
    D = len(fft_shape)
    shape = [1] + list(fft_shape) + [batch]
    t = {"c2c": transform_type.c2c, "r2c": transform_type.r2c, "c2r": transform_type.c2r}[kind]
    d = direction.forward if kind == "r2c" else direction.backward if kind == "c2r" else dirn
    cfg = configuration(D, shape, fp, d, t)
    cfg.set_strides_default(False)
    return cfg


def _get_plan(cfg, adjoint):
    stream = torch.cuda.current_stream().cuda_stream
    key = (cfg.to_string(), adjoint, stream)
    p = _plans.get(key)
    if p is None:
        p = (make_adjoint_plan if adjoint else make_plan)(cfg, stream=stream,
                                                          cache=default_cache())
        _plans[key] = p
    return p


def _fp(dtype):
    if dtype in (torch.float32, torch.complex64):
        return precision.f32
    if dtype in (torch.float64, torch.complex128):
        return precision.f64
    raise TypeError(f"unsupported dtype {dtype}")


def _complex_of(fp):
    return torch.complex64 if fp == precision.f32 else torch.complex128


def _real_of(fp):
    return torch.float32 if fp == precision.f32 else torch.float64


class _Transform(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, cfg, out_shape, out_dtype, in_shape, in_dtype):
        x = x.contiguous()
        y = torch.empty(out_shape, dtype=out_dtype, device=x.device)
        _get_plan(cfg, False).execute(x, y)
        ctx.cfg = cfg
        ctx.in_shape, ctx.in_dtype = in_shape, in_dtype
        return y

    @staticmethod
    def backward(ctx, gy):
        gy = gy.contiguous()
        gx = torch.empty(ctx.in_shape, dtype=ctx.in_dtype, device=gy.device)
        _get_plan(ctx.cfg, True).execute(gy, gx)
        return gx, None, None, None, None, None


def _split(x, ndim):
    if ndim < 1 or ndim > 3 or x.dim() < ndim:
        raise ValueError("ndim must be 1..3 and not exceed the tensor rank")
    fft_rev = list(x.shape[x.dim() - ndim:])  # (Nd, ..., N1)
    batch = math.prod(x.shape[:x.dim() - ndim]) if x.dim() > ndim else 1
    return fft_rev, batch


def fftn(x, ndim=1, direction_=-1):
    """Unnormalised complex transform over the last `ndim` axes (sign `direction_`)."""
    if not x.is_complex():
        raise TypeError("fftn expects a complex tensor")
    fp = _fp(x.dtype)
    fft_rev, batch = _split(x, ndim)
    cfg = _cfg("c2c", fft_rev[::-1], batch, fp, direction(direction_))
    return _Transform.apply(x, cfg, x.shape, x.dtype, x.shape, x.dtype)


def rfftn(x, ndim=1):
    """Unnormalised forward real-to-complex transform; last axis halved to N1/2+1."""
    if x.is_complex():
        raise TypeError("rfftn expects a real tensor")
    fp = _fp(x.dtype)
    fft_rev, batch = _split(x, ndim)
    out_shape = list(x.shape[:-1]) + [x.shape[-1] // 2 + 1]
    cfg = _cfg("r2c", fft_rev[::-1], batch, fp, direction.forward)
    return _Transform.apply(x, cfg, out_shape, _complex_of(fp), x.shape, x.dtype)


def irfftn(X, n1, ndim=1):
    """Unnormalised backward complex-to-real transform; `n1` is the real length of the last
    axis (X's last axis must be n1 // 2 + 1)."""
    if not X.is_complex():
        raise TypeError("irfftn expects a complex tensor")
    if X.shape[-1] != n1 // 2 + 1:
        raise ValueError("last axis must have n1 // 2 + 1 entries")
    fp = _fp(X.dtype)
    fft_rev, batch = _split(X, ndim)
    fft_rev[-1] = n1
    out_shape = list(X.shape[:-1]) + [n1]
    cfg = _cfg("c2r", fft_rev[::-1], batch, fp, direction.backward)
    return _Transform.apply(X, cfg, out_shape, _real_of(fp), X.shape, X.dtype)
