"""The benchmark CLI (reference benchmark interface and CSV columns) on the GPU: every
combination of in-/out-of-place, real/complex and direction runs, checks its result against
torch.fft and prints one CSV row per (M, N)."""
import contextlib
import io

import pytest

from gpu_util import have_gpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not have_gpu(), reason="needs a HIP GPU")]


@pytest.mark.parametrize("args", [
    ["-i", "-m", "1", "-m", "8", "-k", "300", "sc", "16", "60", "256"],
    ["-o", "-v", "-m", "3", "-b", "4M", "dc", "64", "105"],
    ["-i", "-m", "2", "-k", "100", "sr", "32", "27"],
    ["-o", "-v", "-m", "1", "-k", "77", "sr", "128", "45"],
    ["-i", "-v", "-m", "4", "-k", "33", "dr", "48"],
])
def test_benchmark_cli_rows(args):
    from bbfft_amd import benchmark
    out = io.StringIO()
    with contextlib.redirect_stdout(out):
        assert benchmark.main(args + ["--reps", "2"]) == 0
    lines = out.getvalue().strip().splitlines()
    assert lines[0] == ",".join(benchmark.COLUMNS)
    nm = args.count("-m") or 1
    nn = sum(1 for a in args[args.index(next(a for a in args if a in ("sc", "dc", "sr", "dr"))) + 1:])
    assert len(lines) == 1 + nm * nn
    for row in lines[1:]:
        f = row.split(",")
        assert f[0] in "sd" and f[5] in "rc" and float(f[7]) > 0 and float(f[8]) > 0


@pytest.mark.parametrize("args", [
    ["-d", "-k", "40", "16.16", "70.16", "128.512"],
    ["-s", "-k", "25", "1120.32", "3.60"],
])
def test_transpose_fft_transpose(args):
    """Transpose + unit-stride plan + transpose equals the M-batched plan (reference
    examples/transpose_fft_transpose/tft.cpp:148-168 checks the same equality)."""
    from bbfft_amd import tft
    out = io.StringIO()
    with contextlib.redirect_stdout(out):
        assert tft.main(args + ["--reps", "2"]) == 0
    text = out.getvalue()
    assert text.count("Speed-up:") == len(args) - 3
    assert "Error" not in text
