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


def race_check(source, call, in_bytes, out_bytes, grid, block, in_place=False, table=None,
               clang=False):
    """Run a generated kernel once under ThreadSanitizer (see hip_emu.hpp): `call` is the
    kernel call with `in`, `out` (and `table`) in scope, e.g. "k((const float2*)in,
    (float2*)out, 4ul)". Buffers are zero-filled (the kernels' access patterns do not depend on
    the data) and end at an inaccessible page, so accesses past them are reported too; `table`
    (uint64 array) is the ragged job table. Returns (ok, report): ok is False when
    ThreadSanitizer reported a race, an access left the buffers, or the run failed."""
    import hashlib as _h
    os.makedirs(_BUILD, exist_ok=True)
    tab = ""
    if table is not None:
        vals = ",".join(f"{int(v)}ul" for v in table)
        tab = f"static const unsigned long table_data[] = {{{vals}}};\n"
    # Each buffer ends where an inaccessible guard page begins, so an access past its end
    # faults (reported as a bounds error) instead of reading or writing whatever follows.
    tu = ('#include "hip_emu.hpp"\n#include <csignal>\n#include <cstdlib>\n#include <cstring>\n'
          '#include <sys/mman.h>\n#include <unistd.h>\n' + source + "\n" + tab +
          "static void* guarded(std::size_t bytes) {\n"
          "  std::size_t const pg = 4096, span = (bytes + pg - 1) / pg * pg;\n"
          "  char* p = (char*)mmap(nullptr, span + pg, PROT_READ | PROT_WRITE,\n"
          "                        MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);\n"
          "  mprotect(p + span, pg, PROT_NONE);\n"
          "  return p + span - bytes;\n}\n"
          "static void on_fault(int) {\n"
          '  const char m[] = "bounds error: access outside the kernel\'s buffers\\n";\n'
          "  write(2, m, sizeof(m) - 1);\n  _exit(67);\n}\n"
          "int main() {\n"
          "  std::signal(SIGSEGV, on_fault);\n"
          f"  void* in = guarded({max(in_bytes, 16)});\n" +
          ("  void* out = in;\n" if in_place else f"  void* out = guarded({max(out_bytes, 16)});\n") +
          ("  const unsigned long* table = table_data;\n" if table is not None else "") +
          f"  int rc = emu::launch({grid}u, {block}u, [&] {{ {call}; }});\n"
          '  std::printf("launch rc %d\\n", rc);\n'
          "  std::fflush(stdout);\n"
          "  _exit(rc); // races are reported as they happen; skip the sanitizer's exit path\n"
          "}\n")
    with open(os.path.join(_HERE, "hip_emu.hpp"), "rb") as f:
        shim = f.read()
    cxx = CLANG if clang or "BBFFT_HIP_HALF_PRELUDE" in tu else "g++"
    h = _h.sha1(tu.encode() + shim + cxx.encode()).hexdigest()[:16]
    exe = os.path.join(_BUILD, f"race_{h}")
    if not os.path.exists(exe):
        src = os.path.join(_BUILD, f"race_{h}.{os.getpid()}.cpp")
        with open(src, "w") as f:
            f.write(tu)
        tmp = exe + f".tmp{os.getpid()}"
        subprocess.run([cxx, "-O1", "-g", "-std=c++17", "-ffp-contract=off", "-w",
                        "-fsanitize=thread", "-I", _HERE, src, "-o", tmp], check=True)
        os.replace(tmp, exe)
        os.remove(src)
    env = dict(os.environ, TSAN_OPTIONS=os.environ.get(
        "BBFFT_RACE_TSAN_OPTIONS", "halt_on_error=1 exitcode=66 report_signal_unsafe=0 handle_segv=0"))
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    report = r.stdout + r.stderr + f"\n{exe}: exit status {r.returncode}\n"
    ok = r.returncode == 0 and "ThreadSanitizer" not in report
    return ok, report


_tsan = None


def have_tsan():
    """Whether the host compiler builds and runs ThreadSanitizer programs here."""
    global _tsan
    if _tsan is None:
        os.makedirs(_BUILD, exist_ok=True)
        src = os.path.join(_BUILD, f"tsan_probe.{os.getpid()}.cpp")
        exe = src[:-4]
        with open(src, "w") as f:
            f.write("int main() { return 0; }\n")
        try:
            subprocess.run(["g++", "-fsanitize=thread", src, "-o", exe], check=True,
                           capture_output=True)
            _tsan = subprocess.run([exe], capture_output=True).returncode == 0
        except (OSError, subprocess.CalledProcessError):
            _tsan = False
        for p in (src, exe):
            if os.path.exists(p):
                os.remove(p)
    return _tsan
