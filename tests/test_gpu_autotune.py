"""The autotuner end to end on the GPU: search, interleaved confirmation and the table line it
writes, which the library then accepts as a tuning of the same descriptor."""
import json

import pytest

import bbfft_amd as bb
from gpu_util import have_gpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not have_gpu(), reason="needs a HIP GPU")]


def test_autotune_writes_confirmed_entry(tmp_path, capsys):
    from bbfft_amd import autotune
    out = tmp_path / "table.inc"
    # the confirmation runs longer (sustained) rounds than the search
    assert autotune.main(["--rounds", "1", "--reps", "3", "--confirm", "2",
                          "--confirm-reps", "12", "-o", str(out), "scfo8.64*4096"]) == 0
    lines = [ln for ln in out.read_text().splitlines() if autotune._key(ln)]
    assert len(lines) == 1 and autotune._key(lines[0]) == ("f32", "c2c", "64", "8")
    logged = [json.loads(ln) for ln in capsys.readouterr().out.splitlines() if ln.startswith("{")]
    assert sum("confirm" in d for d in logged) >= 2
    fields = lines[0].split("#")[0].split(":")[1].split()
    fac = [int(v) for v in fields[0].split("x")]
    tuning = {"factorization": fac, "Nt": int(fields[1]), "Mb": int(fields[2]),
              "Kb": int(fields[3]), "lds_pad": int(fields[4]), "no_pad": fields[4] == "0",
              "flags": int(fields[5])}
    kc = bb.kernel_configuration(bb.parse_fft_descriptor("scfo8.64*4096"), tuning=tuning)
    assert kc["factorization"] == fac and kc["Nt"] == tuning["Nt"]
