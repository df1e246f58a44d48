"""Benchmark of the double-batched 1-D FFT on one MI355X.

Workload (headline): plain c2c forward, fp32, M=8 x N=256 x K=32768 (descriptor
"scfo8.256*32768"), out-of-place, inputs uniform[-1,1) resident in HBM. A step is one
plan.execute over the whole tensor. Also reported: the sweep N in {16..512, 60, 105, 243, 384}
for fp32 and fp64 (same M, K), device-to-device copy bandwidths (torch's copy and hand-written
streaming-copy kernels: the measured ceiling), the CPU baseline (the oracle's C port of the
reference algorithm on a bounded sample, host cores), the 32 x 128^3 real 3-D transforms
(fused and unfused) and their differentiated filter pipeline (forward vs backward time), the
ragged serving workload (4096 mixed jobs in one graph-replayed launch vs per-job plain
launches) and the N=512 precision variant (fp16 / bf16 storage with the MFMA-split radix-16
stage, errors against an FP64 FFT of the same values).

Accounting (reference benchmark/adapter.hpp:51-52): flops = 5 N log2 N M K, bytes = 2 M N K
sizeof(complex). Timing: HIP events on the plan's stream around the timed steps.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "double-batched-fft-library_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
# HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE (KiB per dispatch), calibrated
# against a 1 GiB device copy on MI355X (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000):
#   rocprofv3 --pmc FETCH_SIZE -- python3 bench.py --no-sweep --no-cpu --steps 3 --warmup 1
#   rocprofv3 --pmc WRITE_SIZE -- python3 bench.py --no-sweep --no-cpu --steps 3 --warmup 1
#   rocprofv3 --pmc FETCH_SIZE|WRITE_SIZE -- python3 bench.py --copy-probe
# Keyed by kernel name; reported only when the timed kernel is the one measured.
MEASURED_TRAFFIC_BYTES = {
    "bbfft_m1_M8_Mb8_N256_f8x8x4_Nt32_Kb1_f32_c2c_is1_8_2048_os1_8_2048_p16_in0_g3":
        (2.0 * 262181.5 + 1.0 * 524288.0) * 1024,
    "bbfft_m1_M8_Mb8_N256_f8x4x8_Nt32_Kb1_f32_c2c_is1_8_2048_os1_8_2048_p16_in0_g67":
        (2.0 * 262182.5 + 1.0 * 524288.0) * 1024,
}
METRIC = "GFLOP/s (5N·log2N) + %HBM peak; MFMA-split vs plain fp32 at equal FP64-err"


# Tensors whose input plus output fit the 256 MB last-level cache (Infinity Cache) several
# times over would be partly served from it between launches; below this footprint the timed
# launches cycle through enough (x, y) pairs to span ROTATE_BYTES, so every launch reads HBM.
ROTATE_BELOW_BYTES = 1 << 30
ROTATE_BYTES = 2 << 30


def buffer_pairs(nbytes_in_out):
    """Number of (x, y) pairs a timing loop cycles through (1 when no rotation is needed)."""
    if nbytes_in_out >= ROTATE_BELOW_BYTES:
        return 1
    return min(64, -(-ROTATE_BYTES // int(nbytes_in_out)))


def time_plan(bb, torch, desc, steps, warmup, cache):
    cfg = bb.parse_fft_descriptor(desc)
    M, N, K = cfg.shape[:3]
    fp = int(cfg.fp)
    rdt = torch.float32 if fp == 4 else torch.float64
    g = torch.Generator(device="cuda")
    g.manual_seed(N)
    nbytes = 2.0 * M * N * K * 2 * fp
    P = buffer_pairs(nbytes)
    xs = [torch.view_as_complex(torch.rand((K, N, M, 2), generator=g, device="cuda", dtype=rdt)
                                * 2 - 1).contiguous() for _ in range(P)]
    ys = [torch.empty_like(xs[0]) for _ in range(P)]
    plan = bb.make_plan(cfg, cache=cache)
    for i in range(warmup):
        plan.execute(xs[i % P], ys[i % P])
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for i in range(steps):
        plan.execute(xs[i % P], ys[i % P])
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / steps
    # per-launch times as the reference harness reports them (min and median of 20
    # individually event-timed launches, after the back-to-back timed region)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(20)]
    for i, (a, b) in enumerate(ev):
        a.record()
        plan.execute(xs[i % P], ys[i % P])
        b.record()
    torch.cuda.synchronize()
    per = sorted(a.elapsed_time(b) for a, b in ev)
    flops = 5.0 * N * math.log2(N) * M * K
    launch = plan.launches()[0]
    del xs, ys
    return {"desc": desc, "ms": ms, "gflops": flops / ms / 1e6, "GBps": nbytes / ms / 1e6,
            "frac": nbytes / ms / 1e6 / HBM_PEAK_GBPS, "bytes": nbytes, "kernel": launch,
            "buffer_pairs": P, "launch_min_ms": per[0],
            "launch_median_ms": 0.5 * (per[9] + per[10])}


def torch_fft_comparator(torch, desc, steps):
    """The same transform through torch.fft (rocFFT / hipFFT on ROCm) on the same tensor
    layout: a vendor-library reference point, not part of the library."""
    import bbfft_amd as bb
    cfg = bb.parse_fft_descriptor(desc)
    M, N, K = cfg.shape[:3]
    if cfg.dim != 1 or int(cfg.type) != 0:
        return None
    rdt = torch.float32 if int(cfg.fp) == 4 else torch.float64
    x = torch.view_as_complex(torch.rand((K, N, M, 2), device="cuda", dtype=rdt) * 2 - 1)
    x = x.contiguous()
    out = torch.empty_like(x)
    fn = (lambda: torch.fft.fft(x, dim=1, out=out)) if int(cfg.dir) < 0 else \
        (lambda: torch.fft.ifft(x, dim=1, norm="forward", out=out))
    for _ in range(5):
        fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(steps):
        fn()
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / steps
    nbytes = 2.0 * x.numel() * x.element_size()
    del x, out
    return {"ms": round(ms, 5), "GBps": round(nbytes / ms / 1e6, 1),
            "GFLOPs": round(5.0 * N * math.log2(N) * M * K / ms / 1e6, 1)}


def stream_copy_bandwidth(torch, nbytes=1 << 30, reps=10):
    """Read + write GB/s of the hand-written gfx950 copy kernels (lib/libbbfft_stream.so):
    the practical streaming ceiling for a 1 GiB -> 1 GiB pass. None if the probe is not built."""
    import ctypes
    so = os.path.join(ROOT, "double-batched-fft-library_amd", "lib", "libbbfft_stream.so")
    if not os.path.exists(so):
        return None
    fn = ctypes.CDLL(so).bbfft_stream_copy
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                   ctypes.c_void_p]
    fn.restype = ctypes.c_int
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").uniform_()
    b = torch.empty_like(a)
    stream = torch.cuda.current_stream().cuda_stream
    out = {}
    for kind, name in ((0, "16B_per_item"), (1, "16KiB_tiles")):
        def run():
            rc = fn(a.data_ptr(), b.data_ptr(), nbytes, kind, stream)
            if rc != 0:
                raise RuntimeError(f"bbfft_stream_copy failed ({rc})")
        run()
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            run()
        e.record()
        torch.cuda.synchronize()
        out[name] = round(2.0 * nbytes * reps / (s.elapsed_time(e) / 1e3) / 1e9, 1)
    assert torch.equal(a, b)
    del a, b
    return out


def copy_bandwidth(torch, nbytes=1 << 30, reps=10):
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").uniform_()
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        b.copy_(a)
    e.record()
    torch.cuda.synchronize()
    return 2.0 * nbytes * reps / (s.elapsed_time(e) / 1e3) / 1e9


def ragged_serving(bb, torch, cache, steps):
    """The seeded mixed serving workload (4096 c2c fp32 jobs, N in {12, 60, 100, 128, 243, 384,
    500, 512}, K log-uniform in [64, 65536], M 1..4, alternating directions) in place on one
    buffer: one persistent launch replayed from a HIP graph, against the same jobs as 4096
    plain plan launches (each job's default plan) on the same stream."""
    from bbfft_amd import ragged
    jobs = ragged.synthetic_serving_jobs()
    n = ragged.total_elements(jobs)
    torch.cuda.empty_cache()
    free, _ = torch.cuda.mem_get_info()
    if free < n * 8 + (8 << 30):
        return {"skipped": f"needs {n * 8 / 1e9:.0f} GB free device memory"}
    buf = torch.empty(n, dtype=torch.complex64, device="cuda")
    g = torch.Generator(device="cuda")
    g.manual_seed(4096)

    def refill():
        torch.rand(torch.view_as_real(buf).shape, generator=g, device="cuda",
                   out=torch.view_as_real(buf))

    def timed(fn, reps):
        refill()
        fn()
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            fn()
        e.record()
        torch.cuda.synchronize()
        return s.elapsed_time(e) / reps

    nbytes = sum(2.0 * j.elements * 8 for j in jobs)
    flops = sum(5.0 * j.N * math.log2(j.N) * j.M * j.K for j in jobs)
    plan = ragged.make_ragged_plan(jobs, 4, cache=cache, graph=True)
    ms = timed(lambda: plan.execute(buf, buf), max(3, min(10, steps // 4)))
    plains = [(bb.make_plan(j.descriptor(), cache=cache), j) for j in jobs]

    def per_job():
        for p, j in plains:
            p.execute(buf[j.in_offset:j.in_offset + j.elements])

    ms_plain = timed(per_job, 2)
    launch = plan.launches()[0]
    del buf
    torch.cuda.empty_cache()
    return {"jobs": len(jobs), "bytes": nbytes, "in_place": True, "graph": True,
            "ms": round(ms, 4), "GFLOPs": round(flops / ms / 1e6, 1),
            "GBps": round(nbytes / ms / 1e6, 1), "frac": round(nbytes / ms / 1e6 / HBM_PEAK_GBPS, 4),
            "per_job_plain_ms": round(ms_plain, 4),
            "per_job_plain_GBps": round(nbytes / ms_plain / 1e6, 1),
            "speedup_vs_per_job": round(ms_plain / ms, 3),
            "kernel": launch["kernel"], "grid": launch["grid"], "block": launch["block"],
            "lds_bytes": launch["lds_bytes"]}


def mfma_split(bb, torch, cache, steps):
    """N=512 c2c forward, M=4 x K=262144: fp16 / bf16 storage through the MFMA-split kernel
    (radix-16 stage as a DFT-matrix GEMM on the matrix cores) and through the vector kernel with
    the same storage, beside the plain fp32 path; each with its error against an FP64 FFT
    (torch.fft on complex128) of the same input values."""
    M, N, K = 4, 512, 262144
    n = M * N * K
    g = torch.Generator(device="cuda")
    g.manual_seed(512)
    re = torch.rand(n, generator=g, device="cuda") * 2 - 1
    im = torch.rand(n, generator=g, device="cuda") * 2 - 1
    cfg = bb.parse_fft_descriptor(f"scfo{M}.{N}*{K}")
    flops = 5.0 * N * math.log2(N) * M * K
    out = {}

    def run(plan, x, y, nbytes, ref):
        plan.execute(x, y)
        torch.cuda.synchronize()
        d = y.to(torch.complex128) - ref
        rel = (torch.linalg.vector_norm(d) / torch.linalg.vector_norm(ref)).item()
        mx = d.abs().max().item()
        del d
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(steps):
            plan.execute(x, y)
        e.record()
        torch.cuda.synchronize()
        ms = s.elapsed_time(e) / steps
        return {"ms": round(ms, 4), "GFLOPs": round(flops / ms / 1e6, 1),
                "GBps": round(nbytes / ms / 1e6, 1),
                "frac": round(nbytes / ms / 1e6 / HBM_PEAK_GBPS, 4),
                "rel_l2_vs_fp64": float(f"{rel:.3e}"), "max_abs_vs_fp64": float(f"{mx:.3e}")}

    for storage, hdt in (("f16", torch.float16), ("bf16", torch.bfloat16)):
        pair = torch.stack([re.to(hdt), im.to(hdt)], dim=1).contiguous()
        x64 = torch.complex(pair[:, 0].double(), pair[:, 1].double())
        ref = torch.fft.fft(x64.view(K, N, M), dim=1).reshape(-1)
        packed = pair.view(torch.int32).reshape(n)
        y = torch.empty(n, dtype=torch.complex64, device="cuda")
        nbytes = n * (4.0 + 8.0)
        for mf in (True, False):
            plan = bb.make_half_input_plan(cfg, storage, mf, cache=cache)
            key = f"{storage}_{'mfma_split' if mf else 'vector'}"
            out[key] = run(plan, packed, y, nbytes, ref)
            out[key]["kernel"] = plan.launches()[0]["kernel"]
        if storage == "f16":
            # plain fp32 storage of the same (f16-exact) values
            x32 = x64.to(torch.complex64)
            plan = bb.make_plan(cfg, cache=cache)
            out["plain_f32"] = run(plan, x32, y, n * 16.0, ref)
            out["plain_f32"]["kernel"] = plan.launches()[0]["kernel"]
            del x32
        del pair, x64, ref, packed, y
        torch.cuda.empty_cache()
    out["speedup_f16_mfma_vs_plain_f32"] = round(out["plain_f32"]["ms"] /
                                                out["f16_mfma_split"]["ms"], 3)
    out["speedup_bf16_mfma_vs_plain_f32"] = round(out["plain_f32"]["ms"] /
                                                 out["bf16_mfma_split"]["ms"], 3)
    return out


def nd_real_3d(bb, torch, cache, steps):
    """3-D real transforms of 32 x 128^3 fp32 (r2c forward and c2r backward, out of place):
    the plan with modes 1 and 2 fused (one kernel per 128 x 65 plane) beside the unfused chain
    of three passes (BBFFT_FUSE=off). Flops 2.5 N log2 N per transform, N = 128^3."""
    n, B = 128, 32
    N = n ** 3
    flops = 2.5 * N * math.log2(N) * B
    nbytes = B * (N * 4.0 + n * n * (n // 2 + 1) * 8.0)
    x = torch.rand(B, n, n, n, device="cuda") * 2 - 1
    X = torch.empty(B, n, n, n // 2 + 1, dtype=torch.complex64, device="cuda")
    out = {}
    for kind, desc, a, b in (("r2c", f"srfo{n}x{n}x{n}*{B}", x, X),
                             ("c2r", f"srbo{n}x{n}x{n}*{B}", X, x)):
        for fused in (True, False):
            os.environ["BBFFT_FUSE"] = "on" if fused else "off"
            plan = bb.make_plan(desc, cache=cache)
            plan.execute(a, b)
            torch.cuda.synchronize()
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record()
            for _ in range(steps):
                plan.execute(a, b)
            e.record()
            torch.cuda.synchronize()
            ms = s.elapsed_time(e) / steps
            out[f"{kind}_{'fused' if fused else 'unfused'}"] = {
                "ms": round(ms, 4), "GFLOPs": round(flops / ms / 1e6, 1),
                "GBps_algorithmic": round(nbytes / ms / 1e6, 1),
                "frac": round(nbytes / ms / 1e6 / HBM_PEAK_GBPS, 4),
                "kernels": [l["kernel"] for l in plan.launches()]}
    os.environ.pop("BBFFT_FUSE", None)
    del x, X
    torch.cuda.empty_cache()
    return out


def autograd_3d_filter(torch, steps):
    """Differentiated 3-D pipeline on 32 x 128^3 fp32 (normal inputs): r2c, multiply by a
    seeded complex 128 x 128 x 65 filter, c2r / 128^3, loss = sum of squares; forward and
    backward (hand-written adjoint kernels through torch.autograd) timed separately with HIP
    events on the current stream. The product with the filter and the 1/128^3 scale run inside
    the c2r's first pass (load callback) and, in the backward, the adjoint's last pass (store
    callback); the same pipeline with the product and scale as separate torch operations is
    timed beside it ("unfused")."""
    from bbfft_amd import autograd as ag
    n, B = 128, 32
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    x = torch.randn(B, n, n, n, generator=g, device="cuda", requires_grad=True)
    f = torch.complex(torch.randn(n, n, n // 2 + 1, generator=g, device="cuda"),
                      torch.randn(n, n, n // 2 + 1, generator=g, device="cuda"))

    def fused():
        z = ag.irfftn(ag.rfftn(x, 3), n, 3, filter=f, scale=1.0 / n ** 3)
        return (z * z).sum()

    def unfused():
        z = ag.irfftn(ag.rfftn(x, 3) * f, n, 3) / n ** 3
        return (z * z).sum()

    def timed(forward, reps):
        for _ in range(2):
            torch.autograd.grad(forward(), x)
        torch.cuda.synchronize()
        tf = tb = 0.0
        for _ in range(reps):
            s, m, e = (torch.cuda.Event(enable_timing=True) for _ in range(3))
            s.record()
            loss = forward()
            m.record()
            torch.autograd.grad(loss, x)
            e.record()
            torch.cuda.synchronize()
            tf += s.elapsed_time(m)
            tb += m.elapsed_time(e)
        return tf / reps, tb / reps

    reps = max(3, min(20, steps // 10))
    tf, tb = timed(fused, reps)
    uf, ub = timed(unfused, reps)
    del x, f
    torch.cuda.empty_cache()
    return {"forward_ms": round(tf, 4), "backward_ms": round(tb, 4),
            "backward_over_forward": round(tb / tf, 3), "reps": reps,
            "unfused": {"forward_ms": round(uf, 4), "backward_ms": round(ub, 4),
                        "backward_over_forward": round(ub / uf, 3)}}


def cpu_baseline(budget_s=15.0):
    """The oracle's C port of the reference algorithm on host cores: the CPU configuration
    (c2c fp32 M=1 x N=64 x K=16384, full K, checked against the FP64 oracle) plus a K=4096
    sample of the headline shape."""
    try:
        import oracle.cpu_port as port
    except Exception:  # noqa: BLE001 - the baseline is optional context
        return None
    res = port.benchmark(budget_s=budget_s)
    res["scipy_pocketfft"] = scipy_comparator(res.get("cores") or os.cpu_count())
    return res


def scipy_comparator(workers, reps=10):
    """Non-reference comparator on the same CPU configuration: scipy.fft (pocketfft) c2c
    forward fp32 over M=1 x N=64 x K=16384 (K rows of 64), min of `reps` runs."""
    try:
        import numpy as np
        import scipy.fft
    except Exception:  # noqa: BLE001 - optional context
        return None
    N, K = 64, 16384
    rng = np.random.default_rng(1)
    x = (rng.uniform(-1, 1, (K, N)) + 1j * rng.uniform(-1, 1, (K, N))).astype(np.complex64)
    scipy.fft.fft(x, axis=1, workers=workers)
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        scipy.fft.fft(x, axis=1, workers=workers)
        best = min(best, time.perf_counter() - t0)
    flops = 5.0 * N * math.log2(N) * K
    return {"GFLOPs": round(flops / best / 1e9, 3), "ms": round(best * 1e3, 4),
            "workers": workers,
            "note": "scipy.fft (pocketfft), not the reference algorithm; labelled comparator"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # ~55 ms of warm-up launches lets the clocks settle before the ~35 ms timed region
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ragged", action="store_true")
    ap.add_argument("--no-mfma", action="store_true")
    ap.add_argument("--no-nd", action="store_true")
    ap.add_argument("--no-autograd", action="store_true")
    ap.add_argument("--desc", default="scfo8.256*32768")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device-to-device copy 5 times (counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}); the double-batched FFT "
              "shards K across GPUs with no exchange, so run it once per GPU.", file=sys.stderr)
        sys.exit(2)

    import torch

    if a.copy_probe:
        gbps = copy_bandwidth(torch, nbytes=1 << 30, reps=5)
        print(json.dumps({"copy_probe_bytes": 1 << 30, "GBps": round(gbps, 1)}), flush=True)
        return

    import bbfft_amd as bb
    cache = bb.default_cache()
    head = time_plan(bb, torch, a.desc, a.steps, a.warmup, cache)
    sweep = []
    if not a.no_sweep:
        for p in "sd":
            for n in (16, 32, 64, 128, 256, 512, 60, 105, 243, 384):
                r = time_plan(bb, torch, f"{p}cfo8.{n}*32768", max(5, a.steps // 2), 2, cache)
                v = torch_fft_comparator(torch, r["desc"], max(5, a.steps // 2))
                sweep.append({"desc": r["desc"], "ms": round(r["ms"], 5),
                              "GFLOPs": round(r["gflops"], 1), "GBps": round(r["GBps"], 1),
                              "frac": round(r["frac"], 4), "buffer_pairs": r["buffer_pairs"],
                              "launch_min_ms": round(r["launch_min_ms"], 5),
                              "launch_median_ms": round(r["launch_median_ms"], 5),
                              "torch_fft_GBps": v["GBps"]})
                print(json.dumps(sweep[-1]), file=sys.stderr, flush=True)
    nd = None
    if not a.no_nd:
        nd = nd_real_3d(bb, torch, cache, a.steps)
        print(json.dumps({"nd_real_3d": nd}), file=sys.stderr, flush=True)
    grad = None
    if not a.no_autograd:
        grad = autograd_3d_filter(torch, a.steps)
        print(json.dumps({"autograd_3d_filter": grad}), file=sys.stderr, flush=True)
    mf = None
    if not a.no_mfma:
        mf = mfma_split(bb, torch, cache, a.steps)
        print(json.dumps({"mfma_split": mf}), file=sys.stderr, flush=True)
    rag = None
    if not a.no_ragged:
        rag = ragged_serving(bb, torch, cache, a.steps)
        print(json.dumps({"ragged": rag}), file=sys.stderr, flush=True)
    copy_gbps = copy_bandwidth(torch)
    stream = stream_copy_bandwidth(torch)
    vendor = torch_fft_comparator(torch, a.desc, min(a.steps, 100))
    cpu = None if a.no_cpu else cpu_baseline()
    cfg = bb.parse_fft_descriptor(a.desc)
    line = {
        "metric": METRIC,
        "value": round(head["gflops"], 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(head["ms"], 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if int(cfg.fp) == 4 else "f64",
        "data": "synthetic",
        "config": {"workload": f"c2c forward {a.desc}: M={cfg.shape[0]} x N={cfg.shape[1]} "
                               f"x K={cfg.shape[2]}, out-of-place, uniform[-1,1) complex",
                   "descriptor": a.desc, "HBM_GBps": round(head["GBps"], 1),
                   "hbm_copy_GBps": round(copy_gbps, 1), "stream_copy_GBps": stream,
                   "frac_of_stream_copy": (round(head["GBps"] / max(stream.values()), 4)
                                           if stream else None),
                   "kernel": head["kernel"]["kernel"],
                   "launch_min_ms": round(head["launch_min_ms"], 6),
                   "launch_median_ms": round(head["launch_median_ms"], 6),
                   "torch_fft_same_tensor": vendor,
                   "sweep": sweep, "nd_real_3d_32x128cubed": nd,
                   "autograd_3d_filter_32x128cubed": grad, "mfma_split_n512": mf,
                   "ragged_serving": rag},
        "roofline": {"bound": "hbm", "achieved": round(head["GBps"], 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(head["frac"], 4),
                     "traffic": MEASURED_TRAFFIC_BYTES.get(head["kernel"]["kernel"])},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
