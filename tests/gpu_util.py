"""Helpers of the GPU tests: run a plan on host data through device buffers and the C ABI."""
import numpy as np

import bbfft_amd as bb

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


def have_gpu():
    return torch is not None and torch.cuda.is_available()


def np_types(fp):
    return (np.float32, np.complex64) if int(fp) == 4 else (np.float64, np.complex128)


def extent(dim, shape, stride):
    return sum((shape[i] - 1) * stride[i] for i in range(dim + 2)) + 1


def io_shapes(cfg):
    """(input shape, output shape) in elements of the respective domain."""
    shape = list(cfg.shape[:cfg.dim + 2])
    half = list(shape)
    half[1] = shape[1] // 2 + 1
    t = int(cfg.type)
    if t == 0:
        return shape, shape
    if t == 1:
        return shape, half
    return half, shape


def execute(cfg, x, inplace=False, plan=None):
    """Run cfg on host array x (flat, input layout); returns the flat host output array."""
    real_t, cplx_t = np_types(cfg.fp)
    t = int(cfg.type)
    in_t = real_t if t == 1 else cplx_t
    out_t = real_t if t == 2 else cplx_t
    ishape, oshape = io_shapes(cfg)
    n_in = extent(cfg.dim, ishape, cfg.istride)
    n_out = extent(cfg.dim, oshape, cfg.ostride)
    x = np.ascontiguousarray(x, dtype=in_t)
    assert x.size >= n_in
    in_bytes = x.nbytes
    out_bytes = n_out * np.dtype(out_t).itemsize
    dev = torch.device("cuda")
    if inplace:
        nbytes = -(-max(in_bytes, out_bytes) // 16) * 16
        src = torch.zeros(nbytes, dtype=torch.uint8, device=dev)
        dst = src
    else:
        src = torch.zeros(-(-in_bytes // 16) * 16, dtype=torch.uint8, device=dev)
        dst = torch.zeros(-(-out_bytes // 16) * 16, dtype=torch.uint8, device=dev)
    src[:in_bytes].copy_(torch.from_numpy(x.view(np.uint8)))
    if plan is None:
        plan = bb.make_plan(cfg, cache=bb.default_cache())
    plan.execute(src.data_ptr(), dst.data_ptr())
    torch.cuda.synchronize()
    return dst[:out_bytes].cpu().numpy().view(out_t).copy()


def approx_equal(a, b, eps):
    """doctest::Approx semantics: |a - b| < eps (1 + max(|a|, |b|))."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) < eps * (1.0 + np.maximum(np.abs(a), np.abs(b)))


def ref_tol(fp, n_total):
    """Reference tolerance 1e2 * eps * sqrt(N) (test/fft.hpp:17-20)."""
    e = float(np.finfo(np.float32 if int(fp) == 4 else np.float64).eps)
    return 1e2 * e * np.sqrt(n_total)


def grid(shape):
    """Index arrays of a column-major tensor of `shape`, flattened consistently."""
    idx = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    return [i.ravel() for i in idx]


def offsets(idx, stride):
    o = np.zeros_like(idx[0])
    for i, s in zip(idx, stride):
        o = o + i * s
    return o
