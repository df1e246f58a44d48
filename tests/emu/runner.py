"""Run generated kernels on the CPU through the fiber emulator (tests/emu/hip_emu.hpp).

Test infrastructure: exercises the generator's indexing, LDS exchange, barriers and real
pre/post passes without a GPU. Compiled kernels are cached under tests/emu/_build/.
"""
import ctypes
import hashlib
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")
CLANG = "/opt/rocm/llvm/bin/clang++"


def _build_so(tu):
    os.makedirs(_BUILD, exist_ok=True)
    with open(os.path.join(_HERE, "hip_emu.hpp"), "rb") as f:
        shim = f.read()  # the shim is part of every translation unit's cache key
    h = hashlib.sha1(tu.encode() + shim).hexdigest()[:16]
    so = os.path.join(_BUILD, f"k_{h}.so")
    if not os.path.exists(so):
        # per-process source and output names: parallel test workers may build the same unit
        src = os.path.join(_BUILD, f"k_{h}.{os.getpid()}.cpp")
        with open(src, "w") as f:
            f.write(tu)
        tmp = so + f".tmp{os.getpid()}"
        # Kernels with half-precision types need clang (_Float16, __bf16, ext_vector_type).
        cxx = CLANG if "BBFFT_HIP_HALF_PRELUDE" in tu else "g++"
        subprocess.run([cxx, "-O1", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-w",
                        "-I", _HERE, src, "-o", tmp], check=True)
        os.replace(tmp, so)
        os.remove(src)
    return ctypes.CDLL(so)


def compile_kernel(source, kernel_name, in_type, out_type):
    """Compile `source` with the emulation shim; returns a ctypes launcher."""
    tu = ('#include "hip_emu.hpp"\n' + source + "\n"
          'extern "C" int emu_launch(const void* in, void* out, unsigned long K, unsigned grid, '
          "unsigned block) {\n"
          f"  return emu::launch(grid, block, [&] {{ {kernel_name}(({in_type}*)in, "
          f"({out_type}*)out, K); }});\n}}\n")
    lib = _build_so(tu)
    fn = lib.emu_launch
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_ulong, ctypes.c_uint, ctypes.c_uint]
    fn.restype = ctypes.c_int
    return fn


def run(bbfft, cfg, x, out_size, grid_cap=None, device_info=None, pass_index=0,
        in_place_buffer=None, tuning=None):
    """Execute 1-D pass `pass_index` of cfg on host arrays through the emulator.

    x: input array (complex or real per the pass); returns out_size output elements. With
    in_place_buffer the kernel reads and writes that buffer (in == out)."""
    info = device_info or bbfft.GFX950_INFO
    kc = bbfft.kernel_configuration(cfg, pass_index, info, tuning=tuning)
    src = bbfft.generate_source(cfg, pass_index, info, tuning=tuning)
    f32 = int(cfg.fp) == 4
    real_t, cplx_t = ("float", "float2") if f32 else ("double", "double2")
    npc, npr = (np.complex64, np.float32) if f32 else (np.complex128, np.float64)
    t = pass_type(cfg, pass_index)
    in_t = real_t if t == 1 else cplx_t
    out_t = real_t if t == 2 else cplx_t
    fn = compile_kernel(src, kc["name"], in_t, out_t)
    M, N, K = kc["M"], kc["N"], kc["K"]
    pairs = t != 0 and N % 2 == 1
    kcols = (K + 1) // 2 if pairs else K
    ntiles = -(-M // kc["Mb"]) * (-(-kcols // kc["Kb"]))
    grid = ntiles if grid_cap is None else min(ntiles, grid_cap)
    if in_place_buffer is not None:
        rc = fn(in_place_buffer.ctypes.data, in_place_buffer.ctypes.data, K, grid,
                kc["work_group_size"])
        if rc != 0:
            raise RuntimeError("divergent barrier in emulated kernel")
        return in_place_buffer
    xin = np.ascontiguousarray(x, dtype=npr if t == 1 else npc)
    out = np.zeros(out_size, dtype=npr if t == 2 else npc)
    rc = fn(xin.ctypes.data, out.ctypes.data, K, grid, kc["work_group_size"])
    if rc != 0:
        raise RuntimeError("divergent barrier in emulated kernel")
    return out


def run_half_input(bbfft, cfg, x_bits, flags, grid_cap=None):
    """Execute a 1-D c2c fp32 kernel reading half-precision (re, im) pairs (x_bits: uint32
    array, real part in the low 16 bits) with fft1d flags `flags`; returns complex64 output."""
    info = bbfft.GFX950_INFO
    tuning = {"flags": flags}
    kc = bbfft.kernel_configuration(cfg, 0, info, tuning=tuning)
    src = bbfft.generate_source(cfg, 0, info, tuning=tuning)
    fn = compile_kernel(src, kc["name"], "unsigned", "float2")
    M, N, K = kc["M"], kc["N"], kc["K"]
    ntiles = -(-M // kc["Mb"]) * (-(-K // kc["Kb"]))
    grid = ntiles if grid_cap is None else min(ntiles, grid_cap)
    xin = np.ascontiguousarray(x_bits, dtype=np.uint32)
    n_out = sum((s - 1) * st for s, st in zip([M, N, K], cfg.ostride[:3])) + 1
    out = np.zeros(n_out, dtype=np.complex64)
    rc = fn(xin.ctypes.data, out.ctypes.data, K, grid, kc["work_group_size"])
    if rc != 0:
        raise RuntimeError("divergent barrier in emulated kernel")
    return out


def run_fused(bbfft, cfg, x, out, grid_cap=None):
    """Execute the fused kernel of a 2-D/3-D configuration from host array x into host array
    out (both in the layouts that kernel reads and writes)."""
    info = bbfft.GFX950_INFO
    fc = bbfft.fused_configuration(cfg, info)
    src = bbfft.generate_fused_source(cfg, info)
    f32 = int(cfg.fp) == 4
    real_t, cplx_t = ("float", "float2") if f32 else ("double", "double2")
    first = fc["first_pass"]
    t = int(cfg.type)
    in_t = real_t if (t == 1 and first == 0) else cplx_t
    out_t = real_t if t == 2 else cplx_t
    fn = compile_kernel(src, fc["name"], in_t, out_t)
    grid = fc["planes"] if grid_cap is None else min(fc["planes"], grid_cap)
    rc = fn(x.ctypes.data, out.ctypes.data, fc["planes"], grid, fc["group_size"])
    if rc != 0:
        raise RuntimeError("divergent barrier in emulated kernel")
    return out


def run_ragged(ragged, jobs, x, out_size, precision=4, grid=7, in_place=False):
    """Execute the persistent ragged kernel for `jobs` on host arrays (`grid` work-groups)."""
    kc = ragged.kernel_configuration(jobs, precision)
    src = ragged.generate_source(jobs, precision)
    table = ragged.job_table(jobs, precision, grid=grid)
    c2 = "float2" if precision == 4 else "double2"
    npc = np.complex64 if precision == 4 else np.complex128
    name = kc["name"]
    tu = ('#include "hip_emu.hpp"\n' + src + "\n"
          'extern "C" int emu_launch_ragged(const void* in, void* out, const unsigned long* t, '
          "unsigned long njobs, unsigned grid, unsigned block) {\n"
          f"  return emu::launch(grid, block, [&] {{ {name}((const {c2}*)in, ({c2}*)out, t, "
          "njobs); });\n}\n")
    fn = _build_so(tu).emu_launch_ragged
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_ulong,
                   ctypes.c_uint, ctypes.c_uint]
    fn.restype = ctypes.c_int
    xin = np.ascontiguousarray(x, dtype=npc)
    out = xin if in_place else np.zeros(out_size, dtype=npc)
    rc = fn(xin.ctypes.data, out.ctypes.data, table.ctypes.data, len(jobs), grid,
            kc["group_size"])
    if rc != 0:
        raise RuntimeError("divergent barrier in emulated kernel")
    return out


def pass_type(cfg, pass_index):
    """Transform type of a 1-D pass: only the real pass (first for r2c, last for c2r) is real."""
    t = int(cfg.type)
    if t == 0 or cfg.dim == 1:
        return t
    if t == 1:
        return 1 if pass_index == 0 else 0
    return 2 if pass_index == cfg.dim - 1 else 0
