"""CPU port of the reference algorithm (oracle/bbfft_cpu.c) -- TEST INFRASTRUCTURE and the
CPU baseline of bench.py. Only tests/ and bench.py's cpu_baseline leg may use it."""
import ctypes
import math
import os
import time

import numpy as np

from . import lib as _lib_loader


def _lib():
    L = _lib_loader()
    if not getattr(L, "_port_ready", False):
        for sfx, T in (("_f32", ctypes.c_float), ("_f64", ctypes.c_double)):
            fn = getattr(L, "cpu_fft_c2c" + sfx)
            fn.argtypes = [ctypes.c_int] * 4 + [ctypes.c_size_t] * 4 + [ctypes.c_void_p] * 2 + [
                ctypes.c_int]
            fn.restype = ctypes.c_int
            info = getattr(L, "cpu_plan_info" + sfx)
            info.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
            info.restype = ctypes.c_int
        L.cpu_factor.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.POINTER(ctypes.c_uint)]
        L.cpu_scramble.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int]
        L.cpu_unscramble.argtypes = L.cpu_scramble.argtypes
        L.cpu_power_of_w.argtypes = [ctypes.c_long, ctypes.c_long,
                                     ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double)]
        L._port_ready = True
    return L


def threads_available():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def fft_c2c(x, M, N, K, direction=-1, precision=4, threads=None):
    """Reference-algorithm c2c of a packed M x N x K tensor (numpy complex, column-major)."""
    ct = np.complex64 if precision == 4 else np.complex128
    xin = np.ascontiguousarray(x, dtype=ct)
    out = np.zeros(M * N * K, dtype=ct)
    fn = getattr(_lib(), "cpu_fft_c2c_f32" if precision == 4 else "cpu_fft_c2c_f64")
    rc = fn(M, N, K, direction, M, M * N, M, M * N, xin.ctypes.data, out.ctypes.data,
            threads or threads_available())
    if rc:
        raise ValueError(f"N={N} is not supported by the port")
    return out


def plan_info(M, N, precision=4):
    out = (ctypes.c_int * 8)()
    fn = getattr(_lib(), "cpu_plan_info_f32" if precision == 4 else "cpu_plan_info_f64")
    fn(M, N, out)
    return {"sbfft": out[0], "factorization": list(out[2:2 + out[1]]), "Mb": out[6]}


def factor(n, index):
    out = (ctypes.c_uint * max(1, index))()
    _lib().cpu_factor(n, index, out)
    return list(out)[:index]


def scramble(fac, index, in0toN=False):
    arr = (ctypes.c_int * len(fac))(*fac)
    return _lib().cpu_scramble(arr, len(fac), index, int(in0toN))


def unscramble(fac, index, in0toN=False):
    arr = (ctypes.c_int * len(fac))(*fac)
    return _lib().cpu_unscramble(arr, len(fac), index, int(in0toN))


def power_of_w(k, n):
    re, im = ctypes.c_double(), ctypes.c_double()
    _lib().cpu_power_of_w(k, n, ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def benchmark(M=8, N=256, budget_s=15.0, precision=4, threads=None):
    """Time the port on a bounded K sample of the M x N x K=32768 workload: min of repeated
    runs (reference benchmark/common.hpp:10-21 takes the min of 10) within ~budget_s."""
    threads = threads or threads_available()
    K = 4096
    rng = np.random.default_rng(0)
    x = (rng.uniform(-1, 1, M * N * K) + 1j * rng.uniform(-1, 1, M * N * K))
    x = x.astype(np.complex64 if precision == 4 else np.complex128)
    fft_c2c(x, M, N, K, -1, precision, threads)  # warm-up
    best = float("inf")
    t_end = time.perf_counter() + budget_s
    reps = 0
    while reps < 10 and (reps < 2 or time.perf_counter() < t_end):
        t0 = time.perf_counter()
        fft_c2c(x, M, N, K, -1, precision, threads)
        best = min(best, time.perf_counter() - t0)
        reps += 1
    flops = 5.0 * N * math.log2(N) * M * K
    return {"value": round(flops / best / 1e9, 3), "unit": "GFLOP/s", "cores": threads,
            "kind": "port",
            "sample": f"c2c forward f{8 * precision} M={M} N={N} K={K} (1/8 of K=32768), "
                      f"min of {reps} runs, {threads} threads"}
