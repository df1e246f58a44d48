"""Oracle for the double-batched FFT -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may import this
package, and only as a checker (or, for the baseline, as the thing timed on the host). The
library never imports it.

Contents
* ``fft`` -- FP64 brute-force DFT over the column-major M x N_1 x ... x N_d x K tensor with
  arbitrary strides (oracle/dft_oracle.c; reference semantics cited there).
* ``uniform_complex`` / ``uniform_real`` -- counter-based (splitmix64) seeded inputs, so the
  same (seed, index) gives the same value on every machine.
* ``tolerance`` / ``rel_l2`` -- the parity bar: relative L2 error <= c * eps * log2(N).

Pinning: tests/test_oracle.py checks ``fft`` against numpy's pocketfft, the reference's
analytic known answers (test/c2c.cpp:28-47, test/r2c.cpp:40-131), Parseval, linearity and the
forward/backward identity. The c2r handling of imag(X[0]) and imag(X[N/2]) on arbitrary
(non-Hermitian) half spectra is pinned by an independent numpy restatement of the reference's
half-length pre-pass and odd-length column pairing (src/base/generator/sbfft_gen.cpp:226-247,
318-351), in 1-D and in the nd order of src/common/algorithm/nd_fft.hpp:89-94.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build():
    """Compile the oracle library (gcc) into oracle/_build/."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        sz5 = ctypes.c_size_t * 5
        _lib.oracle_fft.argtypes = [ctypes.c_uint, sz5, sz5, sz5, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_void_p, ctypes.c_void_p]
        _lib.oracle_fft.restype = ctypes.c_int
    return _lib


def _arr5(v):
    v = list(v) + [0] * (5 - len(v))
    return (ctypes.c_size_t * 5)(*v)


def extent(dim, shape, stride):
    """Number of elements spanned by a tensor with the given shape/strides (max offset + 1)."""
    return sum((shape[i] - 1) * stride[i] for i in range(dim + 2)) + 1


def fft(dim, shape, istride, ostride, type_, direction, x, out_size=None):
    """FP64 oracle. x: complex128 (c2c, c2r) or float64 (r2c) flat array indexed by istride.

    Returns a flat complex128 (c2c, r2c) or float64 (c2r) array indexed by ostride."""
    shape = list(shape)
    oshape = list(shape)
    if type_ == 1:
        oshape[1] = shape[1] // 2 + 1
    n_out = out_size if out_size is not None else extent(dim, oshape if type_ != 2 else shape,
                                                          ostride)
    if type_ == 1:
        xin = np.ascontiguousarray(x, dtype=np.float64)
    else:
        xin = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros(n_out, dtype=np.float64 if type_ == 2 else np.complex128)
    rc = lib().oracle_fft(dim, _arr5(shape), _arr5(istride), _arr5(ostride), type_, direction,
                          xin.ctypes.data, out.ctypes.data)
    if rc != 0:
        raise ValueError("oracle_fft rejected its arguments")
    return out


_M64 = (1 << 64) - 1


def splitmix64(z):
    """Vectorised splitmix64 finaliser over uint64 arrays."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n, offset=0):
    """n values in [-1, 1) keyed by (seed, index)."""
    idx = np.arange(offset, offset + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = idx + np.uint64((seed * 0x632BE59BD9B4E019) & _M64)
    r = splitmix64(key) >> np.uint64(11)
    return r.astype(np.float64) * (2.0 / float(1 << 53)) - 1.0


def uniform_complex(seed, n):
    v = uniform(seed, 2 * n)
    return v[0::2] + 1j * v[1::2]


def uniform_real(seed, n):
    return uniform(seed, n)


def eps(precision):
    return float(np.finfo(np.float32 if precision == 4 else np.float64).eps)


def tolerance(precision, n_total, c=8.0):
    """Relative-L2 bar c * eps * log2(N) (N = product of FFT lengths, at least 2)."""
    return c * eps(precision) * max(1.0, math.log2(max(2, n_total)))


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return float(num / den) if den > 0 else float(num)
