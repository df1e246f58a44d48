"""The host library under AddressSanitizer and UndefinedBehaviorSanitizer: `make sanitize`
builds double-batched-fft-library_amd/_build/san/libbbfft_hip.so (host code instrumented), and
the CPU tests of the host logic, the C ABI, the generator and the emulated kernels run against
it in a child process (BBFFT_LIB points the Python binding at it). Any sanitizer report fails
the test: UBSan does not recover, ASan aborts on its first error."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "double-batched-fft-library_amd")
SAN_LIB = os.path.join(PKG, "_build", "san", "libbbfft_hip.so")

# host logic, C ABI, generator (through the emulated kernels), autotuner host side
SUITES = ["tests/test_host.py", "tests/test_autotune_host.py", "tests/test_oracle.py",
          "tests/test_generator_emu.py", "tests/test_fused_emu.py", "tests/test_ragged_emu.py",
          "tests/test_direct_emu.py", "tests/test_half_emu.py", "tests/test_bench_cli.py"]


def _runtime(name):
    path = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True,
                          text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_host_library_under_asan_and_ubsan():
    asan, stdcxx = _runtime("libasan.so"), _runtime("libstdc++.so.6")
    if not asan or not stdcxx:
        pytest.skip("no AddressSanitizer runtime")
    subprocess.run(["make", "-s", "-j", str(min(8, os.cpu_count() or 1)), "-C", PKG, "sanitize"],
                   check=True, capture_output=True)
    env = dict(os.environ, BBFFT_LIB=SAN_LIB,
               # the sanitizer runtime first; libstdc++ preloaded too, so that its exception
               # interceptors resolve before the library (and its C++ runtime) is loaded
               LD_PRELOAD=f"{asan} {stdcxx}",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        "-m", "not gpu", "-x"] + SUITES,
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=1800)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    assert " passed" in out
