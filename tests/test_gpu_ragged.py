"""Ragged batches on the GPU: one persistent launch (optionally replayed from a HIP graph)
against the FP64 oracle per job and, bitwise, against per-job plain launches of the kernel with
the same factorisation, work-items and M tile (K tile 1); graph replays against eager launches;
the full seeded serving workload (4096 jobs, ~181 GB, in-place) against per-job plain launches.
"""
import numpy as np
import pytest

import bbfft_amd as bb
import oracle
from bbfft_amd import ragged
from gpu_util import have_gpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not have_gpu(), reason="needs a HIP GPU")]

if have_gpu():
    import torch


def _serving_variant_jobs(seed, extra=40, kmax=40):
    """Every (N, M, direction) of the serving workload once (so the kernel is the serving
    kernel from the ahead-of-time archive) plus `extra` random jobs; small K, shuffled."""
    rng = np.random.default_rng(seed)
    jobs = [ragged.Job(M, N, int(rng.integers(1, kmax + 1)), d)
            for N in ragged.SERVING_LENGTHS for M in (1, 2, 3, 4) for d in (-1, 1)]
    for _ in range(extra):
        jobs.append(ragged.Job(int(rng.integers(1, 5)), int(rng.choice(ragged.SERVING_LENGTHS)),
                               int(rng.integers(1, 3 * kmax)), int(rng.choice([-1, 1]))))
    rng.shuffle(jobs)
    return ragged.pack(jobs)


def _plain_plan(job, vc, fp):
    cfg = bb.configuration(1, [job.M, job.N, job.K], bb.precision(fp),
                           bb.direction(job.direction), bb.transform_type.c2c)
    cfg.set_strides_default(False)
    tuning = {"factorization": vc["factorization"], "Nt": vc["Nt"], "Mb": vc["Mb"], "Kb": 1,
              "lds_pad": vc["lds_pad"], "no_pad": vc["lds_pad"] == 0, "flags": 0}
    return bb.make_plan(cfg, cache=bb.default_cache(), tuning=tuning)


def _to_dev(x, fp):
    ct = np.complex64 if fp == 4 else np.complex128
    return torch.from_numpy(np.ascontiguousarray(x, dtype=ct)).to("cuda")


def _check_jobs(jobs, x_host, y_dev, fp, xd=None):
    """Oracle tolerance per job; bitwise equality with the per-job plain launch."""
    y = y_dev.cpu().numpy()
    var = ragged.job_variants(jobs, fp)
    vcs = {}
    xd = _to_dev(x_host, fp) if xd is None else xd
    for j, v in zip(jobs, var):
        seg = x_host[j.in_offset:j.in_offset + j.elements]
        got = y[j.out_offset:j.out_offset + j.elements]
        ref = oracle.fft(1, [j.M, j.N, j.K], [1, j.M, j.M * j.N], [1, j.M, j.M * j.N], 0,
                         j.direction, seg)
        assert oracle.rel_l2(got, ref) <= oracle.tolerance(fp, j.N), j
        if v not in vcs:
            vcs[v] = ragged.variant_configuration(jobs, v, fp)
        out = torch.empty(j.elements, dtype=xd.dtype, device="cuda")
        _plain_plan(j, vcs[v], fp).execute(xd[j.in_offset:j.in_offset + j.elements], out)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), got), ("not bitwise", j)


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("seed", [1, 2])
def test_serving_variants_f32(seed, graph):
    jobs = _serving_variant_jobs(seed)
    kc = ragged.kernel_configuration(jobs)
    assert kc["name"] == ragged.kernel_configuration(ragged.synthetic_serving_jobs())["name"]
    n = ragged.total_elements(jobs)
    x = oracle.uniform_complex(seed, n)
    xd = _to_dev(x, 4)
    yd = torch.zeros_like(xd)
    plan = ragged.make_ragged_plan(jobs, 4, cache=bb.default_cache(), graph=graph)
    (launch,) = plan.launches()
    assert launch["kernel"] == kc["name"] and launch["block"] == kc["group_size"]
    plan.execute(xd, yd)
    torch.cuda.synchronize()
    _check_jobs(jobs, x, yd, 4, xd)


@pytest.mark.parametrize("graph", [False, True])
def test_mixed_f64(graph):
    rng = np.random.default_rng(7)
    jobs = ragged.pack([ragged.Job(int(rng.integers(1, 4)), int(N), int(rng.integers(1, 60)), d)
                        for N, d in [(16, -1), (60, 1), (105, -1), (128, 1), (243, -1),
                                     (17, 1), (512, -1), (60, -1)]])
    n = ragged.total_elements(jobs)
    x = oracle.uniform_complex(8, n)
    xd = _to_dev(x, 8)
    yd = torch.zeros_like(xd)
    plan = ragged.make_ragged_plan(jobs, 8, cache=bb.default_cache(), graph=graph)
    plan.execute(xd, yd)
    torch.cuda.synchronize()
    _check_jobs(jobs, x, yd, 8, xd)


def test_in_place():
    jobs = _serving_variant_jobs(3, extra=10)
    n = ragged.total_elements(jobs)
    x = oracle.uniform_complex(3, n)
    buf = _to_dev(x, 4)
    plan = ragged.make_ragged_plan(jobs, 4, cache=bb.default_cache(), graph=True)
    plan.execute(buf, buf)
    torch.cuda.synchronize()
    ref = torch.zeros_like(buf)
    ragged.make_ragged_plan(jobs, 4, cache=bb.default_cache(), graph=False).execute(
        _to_dev(x, 4), ref)
    torch.cuda.synchronize()
    assert torch.equal(buf, ref)


def test_graph_replays_and_recapture():
    """Replays of the captured launch, with eager work and synchronisations in between, each
    equal to an eager launch on the same input; new buffers re-capture."""
    jobs = _serving_variant_jobs(4, extra=5)
    n = ragged.total_elements(jobs)
    x = _to_dev(oracle.uniform_complex(4, n), 4)
    y = torch.empty_like(x)
    g_plan = ragged.make_ragged_plan(jobs, 4, cache=bb.default_cache(), graph=True)
    e_plan = ragged.make_ragged_plan(jobs, 4, cache=bb.default_cache(), graph=False)
    ref = torch.empty_like(x)
    for rep in range(4):
        x.mul_(0.5 + 0.25 * rep).add_(0.125)  # eager work between replays
        y.fill_(float("nan"))
        torch.cuda.synchronize()
        g_plan.execute(x, y)
        e_plan.execute(x, ref)
        torch.cuda.synchronize()
        assert torch.equal(y, ref), rep
    y2 = torch.full_like(x, float("nan"))
    g_plan.execute(x, y2)
    torch.cuda.synchronize()
    assert torch.equal(y2, ref)
    g_plan.execute(x, y)
    torch.cuda.synchronize()
    assert torch.equal(y, ref)


def test_events_and_stream():
    jobs = _serving_variant_jobs(5, extra=0, kmax=8)
    n = ragged.total_elements(jobs)
    x = _to_dev(oracle.uniform_complex(5, n), 4)
    y = torch.zeros_like(x)
    s = torch.cuda.Stream()
    plan = ragged.make_ragged_plan(jobs, 4, stream=s.cuda_stream, cache=bb.default_cache())
    ev_in, ev_out = torch.cuda.Event(), torch.cuda.Event()
    ev_out.record()  # creates the underlying hipEvent_t
    ev_in.record()
    plan.execute(x, y, signal=ev_out.cuda_event, wait=[ev_in.cuda_event])
    ev_out.synchronize()
    ref = torch.empty_like(y)
    ragged.make_ragged_plan(jobs, 4, cache=bb.default_cache(), graph=False).execute(x, ref)
    torch.cuda.synchronize()
    assert torch.equal(y, ref)


def test_misaligned_pointer_rejected():
    jobs = ragged.pack([ragged.Job(1, 16, 4)])
    x = torch.zeros(200, dtype=torch.uint8, device="cuda")
    plan = ragged.make_ragged_plan(jobs, 4, cache=bb.default_cache())
    with pytest.raises(bb.BbfftError):
        plan.execute(x.data_ptr() + 4, x.data_ptr() + 64)


def test_full_serving_workload_in_place():
    """4096 seeded jobs (N in {12, 60, 100, 128, 243, 384, 500, 512}, K log-uniform in
    [64, 65536], M 1..4, alternating directions) in one graph-replayed launch, in place on one
    ~181 GB buffer; every job compared bitwise with its per-job plain launch on regenerated
    input, and the three smallest jobs with the oracle."""
    jobs = ragged.synthetic_serving_jobs()
    n = ragged.total_elements(jobs)
    torch.cuda.empty_cache()
    free, _ = torch.cuda.mem_get_info()
    need = n * 8 + 2 * max(j.elements for j in jobs) * 8 + (4 << 30)
    if free < need:
        pytest.skip(f"needs {need / 1e9:.0f} GB free device memory, {free / 1e9:.0f} GB free")
    print(f"serving workload: {len(jobs)} jobs, {n * 8 / 1e9:.1f} GB in place", flush=True)

    def job_input(i, j, out):
        g = torch.Generator(device="cuda")
        g.manual_seed(100003 + i)
        torch.rand((j.elements, 2), generator=g, device="cuda", out=out)
        out.mul_(2).sub_(1)

    buf = torch.empty(n, dtype=torch.complex64, device="cuda")
    bview = torch.view_as_real(buf)
    for i, j in enumerate(jobs):
        job_input(i, j, bview[j.in_offset:j.in_offset + j.elements])
    plan = ragged.make_ragged_plan(jobs, 4, cache=bb.default_cache(), graph=True)
    plan.execute(buf, buf)
    torch.cuda.synchronize()
    var = ragged.job_variants(jobs)
    vcs = {v: ragged.variant_configuration(jobs, v) for v in set(var)}
    emax = max(j.elements for j in jobs)
    xin = torch.empty(emax, dtype=torch.complex64, device="cuda")
    yout = torch.empty(emax, dtype=torch.complex64, device="cuda")
    mism = torch.zeros((), dtype=torch.int64, device="cuda")
    for i, (j, v) in enumerate(zip(jobs, var)):
        job_input(i, j, torch.view_as_real(xin[:j.elements]))
        _plain_plan(j, vcs[v], 4).execute(xin[:j.elements], yout[:j.elements])
        mism += (yout[:j.elements] != buf[j.out_offset:j.out_offset + j.elements]).sum()
    torch.cuda.synchronize()
    assert mism.item() == 0
    for i in sorted(range(len(jobs)), key=lambda i: jobs[i].elements)[:3]:
        j = jobs[i]
        job_input(i, j, torch.view_as_real(xin[:j.elements]))
        x = xin[:j.elements].cpu().numpy()
        ref = oracle.fft(1, [j.M, j.N, j.K], [1, j.M, j.M * j.N], [1, j.M, j.M * j.N], 0,
                         j.direction, x)
        got = buf[j.out_offset:j.out_offset + j.elements].cpu().numpy()
        assert oracle.rel_l2(got, ref) <= oracle.tolerance(4, j.N), j
    del buf, xin, yout
    torch.cuda.empty_cache()
