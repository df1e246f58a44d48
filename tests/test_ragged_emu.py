"""Ragged batches on the CPU: the persistent kernel (generated HIP, compiled for the host and
run by the fiber emulator) against the FP64 oracle per job and against the plain kernel of
each job with the same tiling (bitwise); plus the host-side job table and variant logic."""
import numpy as np
import pytest

import bbfft_amd as bb
import oracle
from bbfft_amd import ragged
from emu import runner


def _ref(job, x, fp):
    shape = [job.M, job.N, job.K]
    st = [1, job.M, job.M * job.N]
    seg = x[job.in_offset:job.in_offset + job.elements]
    return oracle.fft(1, shape, st, st, 0, job.direction, seg)


def _mixed_jobs(seed, n, lengths, kmax=9, mmax=4):
    rng = np.random.default_rng(seed)
    jobs = [ragged.Job(M=int(rng.integers(1, mmax + 1)), N=int(rng.choice(lengths)),
                       K=int(rng.integers(1, kmax + 1)), direction=int(rng.choice([-1, 1])))
            for _ in range(n)]
    return ragged.pack(jobs)


@pytest.mark.parametrize("fp", [4, 8])
@pytest.mark.parametrize("seed,lengths", [
    (1, (12, 60, 100, 128)),
    (2, (243, 384, 500, 512)),
    (3, (2, 3, 7, 16, 17, 64)),
])
def test_ragged_vs_oracle(fp, seed, lengths):
    jobs = _mixed_jobs(seed, 12, lengths)
    n = ragged.total_elements(jobs)
    x = oracle.uniform_complex(seed, n)
    y = runner.run_ragged(ragged, jobs, x, n, precision=fp, grid=5)
    for j in jobs:
        ref = _ref(j, x, fp)
        got = y[j.out_offset:j.out_offset + j.elements]
        assert oracle.rel_l2(got, ref) <= oracle.tolerance(fp, j.N), j


def test_ragged_in_place_and_scattered_offsets():
    """In-place execution, and jobs whose outputs are scattered in a different order than
    their inputs (offsets are per job)."""
    jobs = _mixed_jobs(4, 9, (16, 60, 128, 243))
    n = ragged.total_elements(jobs)
    x = oracle.uniform_complex(4, n)
    y = runner.run_ragged(ragged, jobs, x.astype(np.complex64).copy(), n, grid=3,
                          in_place=True)
    for j in jobs:
        assert oracle.rel_l2(y[j.out_offset:j.out_offset + j.elements],
                             _ref(j, x, 4)) <= oracle.tolerance(4, j.N)
    # outputs in reverse job order, with gaps
    off, moved = 7, []
    for j in reversed(jobs):
        moved.append((j, off))
        off += j.elements + 3
    jobs2 = [ragged.Job(j.M, j.N, j.K, j.direction, j.in_offset, o) for j, o in moved[::-1]]
    y2 = runner.run_ragged(ragged, jobs2, x, off, grid=4)
    for j in jobs2:
        assert oracle.rel_l2(y2[j.out_offset:j.out_offset + j.elements],
                             _ref(j, x, 4)) <= oracle.tolerance(4, j.N)


def test_ragged_matches_plain_kernels_bitwise():
    """Each job's result equals the plain kernel with the variant's factorisation, work-items
    and M tile (the K tile does not change the arithmetic)."""
    jobs = _mixed_jobs(5, 10, (12, 60, 100, 128, 243, 384, 500, 512), kmax=5)
    n = ragged.total_elements(jobs)
    x = oracle.uniform_complex(5, n).astype(np.complex64)
    y = runner.run_ragged(ragged, jobs, x, n, grid=6)
    var = ragged.job_variants(jobs)
    for j, v in zip(jobs, var):
        vc = ragged.variant_configuration(jobs, v)
        cfg = bb.configuration(1, [j.M, j.N, j.K], bb.precision.f32, bb.direction(j.direction),
                               bb.transform_type.c2c)
        cfg.set_strides_default(False)
        tuning = {"factorization": vc["factorization"], "Nt": vc["Nt"], "Mb": vc["Mb"],
                  "Kb": 1, "lds_pad": vc["lds_pad"], "no_pad": vc["lds_pad"] == 0, "flags": 0}
        plain = runner.run(bb, cfg, x[j.in_offset:j.in_offset + j.elements], j.elements,
                           tuning=tuning)
        assert np.array_equal(plain, y[j.out_offset:j.out_offset + j.elements]), j


@pytest.mark.parametrize("group,lds", [(512, 81920), (1024, 163840)])
def test_larger_common_group_same_bits(group, lds, monkeypatch):
    """A larger common work-group (K tiles widened to fill it) computes the same bits as the
    default 256-work-item kernel."""
    jobs = _mixed_jobs(6, 8, (12, 60, 243, 500, 512), kmax=6)
    n = ragged.total_elements(jobs)
    x = oracle.uniform_complex(6, n).astype(np.complex64)
    y0 = runner.run_ragged(ragged, jobs, x, n, grid=3)
    monkeypatch.setenv("BBFFT_RAGGED_GROUP", str(group))
    monkeypatch.setenv("BBFFT_RAGGED_LDS", str(lds))
    kc = ragged.kernel_configuration(jobs, 4)
    assert kc["group_size"] == group
    y1 = runner.run_ragged(ragged, jobs, x, n, grid=3)
    assert np.array_equal(y0, y1)


def test_job_table_layout():
    jobs = ragged.pack([ragged.Job(2, 60, 100, -1), ragged.Job(1, 12, 7, 1),
                        ragged.Job(2, 60, 3, -1), ragged.Job(3, 128, 50, 1)])
    kc = ragged.kernel_configuration(jobs)
    var = ragged.job_variants(jobs)
    assert var[0] == var[2] and len(set(var)) == 3 and kc["num_variants"] == 3
    grid = 5
    t = ragged.job_table(jobs, grid=grid)
    nj = len(jobs)
    assert len(t) == 5 * nj + 1 + 2 * grid + 1
    # table order groups jobs by variant, keeping the given order within a variant
    recs = t[nj + 1:5 * nj + 1].reshape(nj, 4)
    order = [next(i for i, j in enumerate(jobs) if j.in_offset == r[0]) for r in recs]
    assert order == sorted(range(nj), key=lambda i: (var[i], i))
    # tile counts: ceil(M / Mb) * ceil(K / Kb) per job, prefix sums
    tiles, weights = [], []
    for p, i in enumerate(order):
        vc = ragged.variant_configuration(jobs, var[i])
        tiles.append(-(-jobs[i].M // vc["Mb"]) * -(-jobs[i].K // vc["Kb"]))
        weights.append(vc["Mb"] * vc["Kb"] * jobs[i].N)
        assert recs[p][1] == jobs[i].out_offset and recs[p][2] == jobs[i].K
        assert recs[p][3] == var[i]
    begin = [0] + list(np.cumsum(tiles))
    assert list(t[:nj + 1]) == begin
    assert kc["total_tiles"] == sum(tiles)
    for v in range(kc["num_variants"]):
        vc = ragged.variant_configuration(jobs, v)
        assert vc["work_group_size"] <= kc["group_size"]
        assert max(vc["factorization"]) <= 16
        assert vc["lds_bytes"] <= kc["lds_bytes"]
    # work split: contiguous, covering, balanced by elements, with the job of each first tile
    first = t[5 * nj + 1:5 * nj + 2 + grid]
    pos = t[5 * nj + 2 + grid:]
    assert first[0] == 0 and first[-1] == sum(tiles) and np.all(np.diff(first.astype(np.int64)) >= 0)
    tile_w = np.repeat(weights, tiles)
    per_wg = [tile_w[first[g]:first[g + 1]].sum() for g in range(grid)]
    ideal = sum(per_wg) / grid  # boundaries are rounded to whole tiles
    assert all(abs(w - ideal) <= max(weights) for w in per_wg)
    for g in range(grid):
        p = int(pos[g])
        assert begin[p] <= first[g] and (first[g] < begin[p + 1] or first[g] == first[g + 1])


@pytest.mark.parametrize("grid", [1, 2, 7, 64])
def test_work_split_grids(grid):
    """Any grid size gives the same results (every tile processed exactly once)."""
    jobs = _mixed_jobs(9, 7, (12, 60, 128, 243), kmax=12)
    n = ragged.total_elements(jobs)
    x = oracle.uniform_complex(9, n)
    y = runner.run_ragged(ragged, jobs, x, n, grid=grid)
    for j in jobs:
        assert oracle.rel_l2(y[j.out_offset:j.out_offset + j.elements],
                             _ref(j, x, 4)) <= oracle.tolerance(4, j.N)


def test_identifier_depends_on_variants_only():
    a = ragged.pack([ragged.Job(2, 60, 100, -1), ragged.Job(1, 128, 7, 1)])
    b = ragged.pack([ragged.Job(1, 128, 9000, 1), ragged.Job(2, 60, 1, -1)])
    c = ragged.pack([ragged.Job(2, 60, 100, 1), ragged.Job(1, 128, 7, 1)])
    name = ragged.kernel_configuration(a)["name"]
    assert name == ragged.kernel_configuration(b)["name"]
    assert name != ragged.kernel_configuration(c)["name"]


@pytest.mark.parametrize("bad", [
    [ragged.Job(0, 16, 4)], [ragged.Job(1, 0, 4)], [ragged.Job(1, 16, 0)],
    [ragged.Job(1, 16, 4, direction=0)], [ragged.Job(1, 4099, 4)], [],
])
def test_invalid_jobs(bad):
    with pytest.raises((bb.BadConfiguration, bb.BbfftError, ValueError)):
        ragged.kernel_configuration(bad)


def test_serving_workload_statistics():
    """The seeded mixed workload: 4096 jobs, N over the eight lengths, K log-uniform in
    [64, 65536], M in 1..4, alternating directions, packed offsets."""
    jobs = ragged.synthetic_serving_jobs()
    assert len(jobs) == 4096
    assert {j.N for j in jobs} == set(ragged.SERVING_LENGTHS)
    assert {j.M for j in jobs} == {1, 2, 3, 4}
    assert all(64 <= j.K <= 65536 for j in jobs)
    assert [j.direction for j in jobs[:4]] == [-1, 1, -1, 1]
    lk = np.log([j.K for j in jobs])
    assert abs(lk.mean() - (np.log(64) + np.log(65536)) / 2) < 0.15
    assert jobs[1].in_offset == jobs[0].elements
    assert ragged.synthetic_serving_jobs() == jobs
    assert ragged.kernel_configuration(jobs)["num_variants"] == 64
