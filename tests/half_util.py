"""Half-precision (re, im) pair storage for the half-input kernels: quantisation of complex
values to IEEE binary16 / bfloat16 pairs packed in uint32 (real part in the low 16 bits), and
back to complex128 (exact)."""
import numpy as np


def _bf16_bits(x):
    """Round float32 values to bfloat16 (nearest, ties to even; finite inputs)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = u + 0x7FFF + ((u >> 16) & 1)
    return (u >> 16).astype(np.uint32)


def pack(z, storage):
    """complex array -> uint32 array of (re, im) half pairs."""
    z = np.asarray(z)
    if storage == "f16":
        re = np.real(z).astype(np.float16).view(np.uint16).astype(np.uint32)
        im = np.imag(z).astype(np.float16).view(np.uint16).astype(np.uint32)
    else:
        re = _bf16_bits(np.real(z).astype(np.float32))
        im = _bf16_bits(np.imag(z).astype(np.float32))
    return (re | (im << 16)).astype(np.uint32)


def unpack(bits, storage):
    """uint32 array of (re, im) half pairs -> complex128 (exact)."""
    bits = np.asarray(bits, dtype=np.uint32)
    lo = (bits & 0xFFFF).astype(np.uint16)
    hi = (bits >> 16).astype(np.uint16)
    if storage == "f16":
        re = lo.view(np.float16).astype(np.float64)
        im = hi.view(np.float16).astype(np.float64)
    else:
        re = (lo.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        im = (hi.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    return re + 1j * im


def heavy_tail(rng, n):
    """Complex values with log-uniform magnitudes over [1e-3, 1e3] and uniform phase."""
    mag = 10.0 ** rng.uniform(-3.0, 3.0, n)
    ph = rng.uniform(0.0, 2.0 * np.pi, n)
    return mag * np.exp(1j * ph)
