"""The C++ drop-in API on the GPU: tests/cpp/cxx_gpu_test.cpp (built by
__graft_entry__.build() against the in-tree library) plans and executes c2c and r2c transforms
through bbfft::make_plan / plan::execute, out of place, in place and with caller-owned HIP
events, addressing its tensors (padded strides included) with bbfft::tensor_indexer, and checks
the reference tests' known answers."""
import os
import subprocess

import pytest

from gpu_util import have_gpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not have_gpu(), reason="needs a HIP GPU")]

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "cxx_gpu_test")


def test_cxx_api_on_gpu():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is missing: build it with __graft_entry__.build()")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("PASSED")
    assert r.stdout.count(" ok") == 7
