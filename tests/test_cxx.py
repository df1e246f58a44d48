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


def test_aot_archive_embedded_in_executable(tmp_path):
    """`python -m bbfft_amd.aot --embed` turns an archive into C++; linked into a program, the
    archive it holds lists the same kernels (bbfft_aot_archive_kernels, host only)."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "double-batched-fft-library_amd"))
    import bbfft_amd as bb
    from bbfft_amd import aot
    arch = str(tmp_path / "k.bin")
    n = bb.aot_generate_shard(["scfo8.16*100", "srfo2.12*5"], arch)
    assert n >= 2
    cpp = str(tmp_path / "embedded.cpp")
    aot.embed([arch], cpp, "my_kernels")
    main = tmp_path / "main.cpp"
    main.write_text('''#include "bbfft_hip.h"
#include <cstdint>
#include <cstdio>
#include <vector>
extern const std::uint8_t my_kernels[];
extern const std::size_t my_kernels_size;
int main() {
    size_t need = 0;
    if (bbfft_aot_archive_kernels(my_kernels, my_kernels_size, nullptr, 0, &need)) return 1;
    std::vector<char> buf(need);
    if (bbfft_aot_archive_kernels(my_kernels, my_kernels_size, buf.data(), need, &need)) return 2;
    std::fputs(buf.data(), stdout);
    return 0;
}
''')
    exe = str(tmp_path / "prog")
    subprocess.run(["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I", INC,
                    "-I", "/opt/rocm/include", str(main), cpp, "-o", exe, "-L", LIBDIR,
                    "-lbbfft_hip", "-Wl,-rpath," + LIBDIR, "-L/opt/rocm/lib", "-lamdhip64",
                    "-Wl,-rpath,/opt/rocm/lib"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    names = {ln.split()[1] for ln in r.stdout.splitlines()}
    assert all(ln.split()[0] == "gfx950" for ln in r.stdout.splitlines())
    expect = {bb.kernel_configuration(bb.parse_fft_descriptor(d))["name"]
              for d in ["scfo8.16*100", "srfo2.12*5"]}
    assert expect <= names and len(names) == n
