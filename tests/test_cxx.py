"""C++ header tests (CPU): tensor_indexer behaviour and that the drop-in C++ API compiles and
links against libbbfft_hip.so."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "double-batched-fft-library_amd", "lib")


def test_tensor_indexer(tmp_path):
    exe = str(tmp_path / "ti")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", INC,
                    os.path.join(ROOT, "tests", "cpp", "tensor_indexer_test.cpp"), "-o", exe],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout


def test_cxx_api_compiles_and_links(tmp_path):
    so = str(tmp_path / "libapi.so")
    subprocess.run(["g++", "-std=c++17", "-O1", "-fPIC", "-shared", "-D__HIP_PLATFORM_AMD__",
                    "-I", INC, "-I", "/opt/rocm/include",
                    os.path.join(ROOT, "tests", "cpp", "cxx_api_compile_test.cpp"), "-o", so,
                    "-L", LIBDIR, "-lbbfft_hip", "-Wl,--no-undefined",
                    "-L/opt/rocm/lib", "-lamdhip64"], check=True)
    assert os.path.getsize(so) > 0
