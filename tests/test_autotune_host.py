"""Host logic of the autotuner: candidate factorisations, table lines and table merging, and
that the compiled-in table parses into valid kernel configurations."""
import math
import os
import re

import pytest

import bbfft_amd as bb
from bbfft_amd import autotune

TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                     "double-batched-fft-library_amd", "tuning", "gfx950.inc")


@pytest.mark.parametrize("n", [16, 60, 64, 105, 128, 243, 256, 384, 512])
def test_candidates_multiply_to_n(n):
    c = autotune.candidate_factorizations(n, 4)
    assert c and all(math.prod(f) == n for f in c)
    assert all(2 <= r <= 32 for f in c for r in f)
    assert len({tuple(f) for f in c}) == len(c)
    assert all(1 <= len(f) <= 4 for f in c)
    assert [len(f) for f in c] == sorted(len(f) for f in c)  # fewest stages first


def test_candidates_include_tuned_orders():
    """Stage orders matter: the table's measured winners are among the candidates."""
    assert [8, 8, 4] in autotune.candidate_factorizations(256, 4)
    assert [16, 4, 8] in autotune.candidate_factorizations(512, 4)
    assert [4, 4, 4, 4] in autotune.candidate_factorizations(256, 4)


def _table_lines():
    with open(TABLE) as f:
        body = f.read().split('R"TABLE(', 1)[1].split(')TABLE"', 1)[0]
    return [ln for ln in body.splitlines() if autotune._key(ln)]


def test_table_entries_are_valid_configurations():
    for ln in _table_lines():
        prec, typ, N, M = autotune._key(ln)
        fields = ln.split("#")[0].split(":")[1].split()
        fac = [int(v) for v in fields[0].split("x")]
        Nt, Mb, Kb, pad, flags = (int(v) for v in fields[1:6])
        base, *storage = typ.split("-")
        nfft = int(N) // 2 if base != "c2c" and int(N) % 2 == 0 else int(N)
        assert math.prod(fac) == nfft, ln
        # access hints and the load-first tile loop; storage bits come from the type
        assert flags & ~(3 | bb.PREFETCH) == 0, ln
        assert not flags & bb.PREFETCH or typ in ("c2c", "r2c"), ln
        if storage == ["ragged"]:  # variants of the persistent ragged kernel
            assert base == "c2c" and prec == "f32" and not flags & bb.PREFETCH, ln
            assert Mb * Nt <= autotune.RAGGED_GROUP, ln
            continue
        if storage:  # half-input kernels: c2c fp32, MFMA entries start with radix 16
            assert base == "c2c" and prec == "f32" and set(storage) <= {"f16", "bf16", "mfma"}
            assert "mfma" not in storage or fac[0] == 16, ln
            continue
        typ = base
        desc = f"{'s' if prec == 'f32' else 'd'}{'c' if typ == 'c2c' else 'r'}" \
               f"{'b' if typ == 'c2r' else 'f'}o{M}.{N}*1000"
        kc = bb.kernel_configuration(bb.parse_fft_descriptor(desc))
        assert kc["factorization"] == fac and kc["Nt"] == Nt, ln
        assert kc["Mb"] == min(Mb, int(M)) and kc["lds_pad"] == pad, ln


def test_merge_and_line_format(tmp_path):
    lines = _table_lines()
    p = tmp_path / "t.inc"
    autotune.write_table(str(p), lines)
    new = "f32 c2c 256 8 : 4x4x4x4 64 8 1 16 3   # 1 -> 2 GB/s, scfo8.256*32768"
    merged = autotune.merge_table(str(p), [new])
    assert len(merged) == len(lines) and merged[-1] == new
    assert sum(autotune._key(ln) == ("f32", "c2c", "256", "8") for ln in merged) == 1
    cfg = bb.parse_fft_descriptor("scfo8.256*32768")
    t = {"factorization": [8, 8, 4], "Nt": 32, "Mb": 8, "Kb": 1, "lds_pad": 16, "flags": 3}
    line = autotune.table_line(cfg, t, 1.0, 0.5, 1e9, "scfo8.256*32768")
    assert re.match(r"f32 c2c 256 8 : 8x8x4 32 8 1 16 3 +# 1000 -> 2000 GB/s", line)
    # the load-first tile loop flag and the occupancy columns survive; storage bits move to
    # the type
    t = dict(t, flags=bb.PREFETCH | 3 | bb.INPUT_F16, waves_per_eu=0, groups_per_cu=2)
    line = autotune.table_line(cfg, t, 1.0, 0.5, 1e9, "scfo8.256*32768")
    assert re.match(r"f32 c2c-f16 256 8 : 8x8x4 32 8 1 16 67 0 2 +# 1000 -> 2000 GB/s", line)


def test_benchmark_cli_k_selection():
    """K fixed (-k) or from the input tensor's memory footprint (-b), reference
    benchmark/test.hpp:9-34, and the byte-size suffixes of benchmark/args.cpp."""
    from bbfft_amd import benchmark
    assert benchmark.parse_bytes("1G") == 1 << 30 and benchmark.parse_bytes("4k") == 4096
    assert benchmark.parse_bytes("123") == 123

    class A:
        k, b = 0, 1 << 30
    assert benchmark.k_for(A, 8, 256, 4, True) == (1 << 30) // (8 * 256 * 8)
    assert benchmark.k_for(A, 1, 1 << 28, 8, True) == 1
    A.k = 77
    assert benchmark.k_for(A, 8, 256, 4, True) == 77


def test_ragged_entries_configure_the_serving_variants():
    """configure_ragged takes the factorisation, work-items, M tile and padding of the table's
    c2c-ragged entry for each (N, M) of the serving workload."""
    from bbfft_amd import ragged
    entries = {}
    for ln in _table_lines():
        prec, typ, N, M = autotune._key(ln)
        if typ == "c2c-ragged":
            f = ln.split("#")[0].split(":")[1].split()
            entries[(int(N), int(M))] = ([int(v) for v in f[0].split("x")], int(f[1]),
                                         int(f[2]), int(f[4]))
    jobs = ragged.synthetic_serving_jobs()
    var = ragged.job_variants(jobs)
    seen = set()
    for j, v in zip(jobs, var):
        if (j.N, j.M) in seen:
            continue
        seen.add((j.N, j.M))
        assert (j.N, j.M) in entries, (j.N, j.M)
        fac, nt, mb, pad = entries[(j.N, j.M)]
        vc = ragged.variant_configuration(jobs, v)
        assert (vc["factorization"], vc["Nt"], vc["Mb"], vc["lds_pad"]) == (fac, nt, mb, pad)
        assert vc["lds_bytes"] <= 32 * 1024
    assert len(seen) == 32
