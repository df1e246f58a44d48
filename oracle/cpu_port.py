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


def cpu_model():
    """The host CPU's model string (/proc/cpuinfo) or the platform's processor name."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:  # pragma: no cover
        pass
    import platform
    return platform.processor() or "unknown"


_native = None


def native_lib():
    """The port compiled with gcc -O3 -march=native -ffp-contract=fast on the host that runs it
    (the CPU baseline's build; the prebuilt oracle library is compiled portably because it is
    built on one machine and run on another). Falls back to the prebuilt library, returning
    (library, build description)."""
    global _native
    if _native is not None:
        return _native
    import subprocess
    import tempfile
    here = os.path.dirname(os.path.abspath(__file__))
    flags = ["-O3", "-march=native", "-ffp-contract=fast"]
    out_dir = tempfile.mkdtemp(prefix="bbfft_port_")
    so = os.path.join(out_dir, "libport_native.so")
    try:
        subprocess.run(["gcc", *flags, "-std=gnu11", "-fPIC", "-shared", "-o", so,
                        os.path.join(here, "bbfft_cpu.c"), "-lm", "-lpthread"],
                       check=True, capture_output=True, timeout=120)
        L = ctypes.CDLL(so)
        desc = "gcc " + " ".join(flags) + " (compiled on this host)"
    except Exception as e:  # noqa: BLE001 - the portable build still measures the port
        L = _lib_loader()
        desc = f"prebuilt oracle library (-O3 -mavx2 -mfma); native build failed: {e!r:.80}"
    for sfx in ("_f32", "_f64"):
        fn = getattr(L, "cpu_fft_c2c" + sfx)
        fn.argtypes = [ctypes.c_int] * 4 + [ctypes.c_size_t] * 4 + [ctypes.c_void_p] * 2 + [
            ctypes.c_int]
        fn.restype = ctypes.c_int
    _native = (L, desc)
    return _native


def _time_port(fn, x, out, M, N, K, threads, budget_s, min_reps=10):
    """Min over up to `min_reps` runs after one warm-up (reference benchmark/common.hpp:10-21),
    stopping early only when the budget is spent after at least two runs."""
    import time as _t
    args = (M, N, K, -1, M, M * N, M, M * N, x.ctypes.data, out.ctypes.data, threads)
    if fn(*args):
        raise ValueError(f"N={N} is not supported by the port")
    best, reps = float("inf"), 0
    t_end = _t.perf_counter() + budget_s
    while reps < min_reps and (reps < 2 or _t.perf_counter() < t_end):
        t0 = _t.perf_counter()
        fn(*args)
        best = min(best, _t.perf_counter() - t0)
        reps += 1
    return best, reps


def benchmark(budget_s=15.0, threads=None):
    """CPU baseline: the port of the reference algorithm on host cores.

    Primary (value): the CPU configuration itself -- c2c forward fp32, M=1 x N=64 x K=16384,
    out of place, inputs uniform[-1,1) (seeded), full K, min of 10 runs -- with its output
    checked against the FP64 DFT oracle. Secondary: a K=4096 sample of the GPU headline shape
    M=8 x N=256 (1/8 of K=32768)."""
    from . import fft as oracle_fft, rel_l2, uniform_complex
    threads = threads or threads_available()
    L, build = native_lib()
    M, N, K = 1, 64, 16384
    x = uniform_complex(1, M * N * K).astype(np.complex64)
    out = np.zeros_like(x)
    best, reps = _time_port(L.cpu_fft_c2c_f32, x, out, M, N, K, threads, budget_s / 2)
    st = [1, M, M * N, 0, 0]
    ref = oracle_fft(1, [M, N, K], st, st, 0, -1, x.astype(np.complex128))
    err = rel_l2(out, ref)
    flops = 5.0 * N * math.log2(N) * M * K
    res = {"value": round(flops / best / 1e9, 3), "unit": "GFLOP/s", "cores": threads,
           "kind": "port",
           "sample": f"c2c forward f32 M={M} N={N} K={K} (the CPU configuration, full K), "
                     f"min of {reps} runs, {threads} threads",
           "ms": round(best * 1e3, 4), "rel_l2_vs_fp64_oracle": float(f"{err:.3e}"),
           "build": build, "cpu_model": cpu_model(), "hardware_concurrency": os.cpu_count()}
    M2, N2, K2 = 8, 256, 4096
    rng = np.random.default_rng(0)
    x2 = (rng.uniform(-1, 1, M2 * N2 * K2) + 1j * rng.uniform(-1, 1, M2 * N2 * K2)).astype(
        np.complex64)
    out2 = np.zeros_like(x2)
    best2, reps2 = _time_port(L.cpu_fft_c2c_f32, x2, out2, M2, N2, K2, threads, budget_s / 2)
    res["headline_shape_sample"] = {
        "GFLOPs": round(5.0 * N2 * math.log2(N2) * M2 * K2 / best2 / 1e9, 3),
        "ms": round(best2 * 1e3, 3),
        "sample": f"c2c forward f32 M={M2} N={N2} K={K2} (1/8 of K=32768), min of {reps2} runs"}
    return res
