"""Differentiable transforms for PyTorch over the gfx950 kernels (thin binding over the C ABI).

    X = autograd.rfftn(x, ndim=3)          # real (..., N3, N2, N1) -> complex (..., N3, N2, N1/2+1)
    x = autograd.irfftn(X, n1=128, ndim=3) # complex half spectrum -> real, unnormalised
    y = autograd.irfftn(X, 128, 3, filter=H, scale=128 ** -3)  # irfftn(X * H) * scale, fused
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
import collections
import math

import torch

from . import configuration, default_cache, direction, kernel_language, make_adjoint_plan
from . import make_plan, precision, transform_type, user_module

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


def irfftn(X, n1, ndim=1, filter=None, scale=None):
    """Unnormalised backward complex-to-real transform; `n1` is the real length of the last
    axis (X's last axis must be n1 // 2 + 1).

    With `filter` (a complex tensor of X's trailing `ndim` axes, applied to every batch) and/or
    `scale` (a number) it computes irfftn(X * filter) * scale, the product taken by the
    transform's first pass as it loads X (a load callback) instead of a separate pass over the
    spectrum; its backward multiplies by conj(filter) * scale as the adjoint's last pass stores
    (a store callback). The filter is a constant: it gets no gradient. Its buffer's address is
    compiled into the kernels (values are read at each call), so keep one contiguous filter
    tensor across calls: a new buffer (or a non-contiguous filter, which is copied) compiles
    new kernels."""
    if not X.is_complex():
        raise TypeError("irfftn expects a complex tensor")
    if X.shape[-1] != n1 // 2 + 1:
        raise ValueError("last axis must have n1 // 2 + 1 entries")
    fp = _fp(X.dtype)
    fft_rev, batch = _split(X, ndim)
    fft_rev[-1] = n1
    out_shape = list(X.shape[:-1]) + [n1]
    cfg = _cfg("c2r", fft_rev[::-1], batch, fp, direction.backward)
    if filter is None and scale is None:
        return _Transform.apply(X, cfg, out_shape, _real_of(fp), X.shape, X.dtype)
    if filter is not None:
        if filter.dtype != X.dtype or filter.device != X.device:
            raise TypeError("the filter must have X's dtype and device")
        if tuple(filter.shape) != tuple(X.shape[X.dim() - ndim:]):
            raise ValueError("the filter must have the shape of X's last ndim axes")
        if filter.requires_grad:
            raise ValueError("the filter is a constant (it gets no gradient)")
        filter = filter.contiguous()
    return _FilteredC2R.apply(X, filter, 1.0 if scale is None else float(scale), cfg,
                              out_shape)


# Callback source of the fused spectral product. The filter's device address and the scale are
# literals of the source (one kernel per filter buffer and scale, cached like any kernel);
# offsets index the spectrum, whose trailing modes repeat every `period` elements. The
# remainder is taken in 32-bit arithmetic when the spectrum has fewer than 2^32 elements (a
# 64-bit remainder per element costs the 16-element work-items enough registers to spill).
# The store callback writes with non-temporal stores, as the library does for kernels with a
# load callback: the output then does not push the filter out of the last-level cache (32 x
# 128^3 backward 0.765-0.770 ms against 0.773 with plain stores).
_FILTER_SOURCE = """
__device__ {t2} bbfft_filter_load({t2}* in, size_t offset) {{
    const {t2}* h = (const {t2}*){addr:#x}ul;
    const {t2} a = in[offset];
    {t2} w;
    w.x = 1;
    w.y = 0;
    if ({has_filter}) w = h[{index}];
    {t2} r;
    r.x = (a.x * w.x - a.y * w.y) * ({t}){scale};
    r.y = (a.x * w.y + a.y * w.x) * ({t}){scale};
    return r;
}}
__device__ void bbfft_filter_store({t2}* out, size_t offset, {t2} v) {{
    const {t2}* h = (const {t2}*){addr:#x}ul;
    {t2} w;
    w.x = 1;
    w.y = 0;
    if ({has_filter}) w = h[{index}];
    {t2} r;
    r.x = (v.x * w.x + v.y * w.y) * ({t}){scale};
    r.y = (v.y * w.x - v.x * w.y) * ({t}){scale};
    __builtin_nontemporal_store(r.x, &out[offset].x);
    __builtin_nontemporal_store(r.y, &out[offset].y);
}}
"""


_FILTER_PLANS_MAX = 16
_filter_plans = collections.OrderedDict()  # least recently used first


def _filter_plan(cfg, filt, scale, adjoint):
    """c2r plan loading X * filt * scale, or its adjoint storing (.) * conj(filt) * scale.
    Plans are kept per filter buffer (its address is compiled into the kernels): a filter that
    stays in one buffer (updated in place, if at all) compiles once; the last
    _FILTER_PLANS_MAX plans are kept."""
    stream = torch.cuda.current_stream().cuda_stream
    addr = filt.data_ptr() if filt is not None else 0
    period = filt.numel() if filt is not None else 1
    key = (cfg.to_string(), adjoint, stream, addr, period, scale)
    p = _filter_plans.get(key)
    if p is not None:
        _filter_plans.move_to_end(key)
    else:
        f32 = cfg.fp == precision.f32
        elems = math.prod(int(v) for v in cfg.shape[:cfg.dim + 2] if v)
        elems = elems // int(cfg.shape[1]) * (int(cfg.shape[1]) // 2 + 1)  # complex side
        index = (f"(unsigned)offset % {period}u" if elems < (1 << 32)
                 else f"offset % {period}ul")
        src = _FILTER_SOURCE.format(t="float" if f32 else "double",
                                    t2="float2" if f32 else "double2", addr=addr,
                                    index=index, scale=float(scale).hex(),
                                    has_filter="true" if filt is not None else "false")
        c = configuration(cfg.dim, list(cfg.shape), cfg.fp, cfg.dir, cfg.type,
                          list(cfg.istride), list(cfg.ostride),
                          user_module(src, None if adjoint else "bbfft_filter_load",
                                      "bbfft_filter_store" if adjoint else None,
                                      kernel_language.hip))
        p = (make_adjoint_plan if adjoint else make_plan)(c, stream=stream,
                                                          cache=default_cache())
        _filter_plans[key] = p
        while len(_filter_plans) > _FILTER_PLANS_MAX:
            _filter_plans.popitem(last=False)
    return p


class _FilteredC2R(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X, filt, scale, cfg, out_shape):
        X = X.contiguous()
        y = torch.empty(out_shape, dtype=_real_of(cfg.fp), device=X.device)
        _filter_plan(cfg, filt, scale, False).execute(X, y)
        ctx.cfg, ctx.scale = cfg, scale
        ctx.in_shape, ctx.in_dtype = X.shape, X.dtype
        ctx.filt = filt  # kept alive: its address is part of the kernels
        return y

    @staticmethod
    def backward(ctx, gy):
        gy = gy.contiguous()
        gX = torch.empty(ctx.in_shape, dtype=ctx.in_dtype, device=gy.device)
        _filter_plan(ctx.cfg, ctx.filt, ctx.scale, True).execute(gy, gX)
        return gX, None, None, None, None
