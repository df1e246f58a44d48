"""Benchmark CLI with the reference benchmark's interface and CSV columns.

    python -m bbfft_amd.benchmark [-i|-o] [-v] [-m M]... [-k K | -b BYTES] (s|d)(r|c) N1 N2 ...

Reference: benchmark/args.cpp:11-100 (options), benchmark/test.hpp:9-34 (K fixed or from the
input tensor's memory footprint), benchmark/result.hpp:12-30 (CSV columns p, N, inplace, M, K,
domain, inverse, time, bw, flops), benchmark/adapter.hpp:21-52 (bytes and flop counts),
benchmark/common.hpp:10-21 (minimum over repetitions). For every M and N a 1-D plan over the
second mode of an M x N x K tensor runs on the current HIP device; each result is checked
against torch.fft before it is timed (HIP events, minimum of `--reps` runs after a warm-up).
"""
import argparse
import math
import sys

COLUMNS = ["p", "N", "inplace", "M", "K", "domain", "inverse", "time", "bw", "flops"]


def parse_bytes(s):
    mult = {"k": 1024, "m": 1024 ** 2, "g": 1024 ** 3}.get(s[-1].lower(), 1)
    return int(s[:-1] if mult != 1 else s) * mult


def k_for(a, M, N, real_bytes, complex_domain):
    """K fixed (-k) or from the input tensor's memory footprint (-b)."""
    if a.k:
        return a.k
    size = real_bytes * (2 if complex_domain else 1)
    return max(1, a.b // (M * N * size))


def run_case(bb, torch, p, d, M, N, K, inplace, inverse, reps):
    fp = 4 if p == "s" else 8
    real = d == "r"
    rdt = torch.float32 if fp == 4 else torch.float64
    cdt = torch.complex64 if fp == 4 else torch.complex128
    kind, dirn = ("r" if real else "c"), ("b" if inverse else "f")
    desc = f"{p}{kind}{dirn}{'i' if inplace else 'o'}{M}.{N}*{K}"
    cfg = bb.parse_fft_descriptor(desc)
    plan = bb.make_plan(cfg, cache=bb.default_cache())
    g = torch.Generator(device="cuda")
    g.manual_seed(N * 131 + M)
    H = N // 2 + 1
    # tensors in bbfft's column-major layout = torch row-major (K, N, M)
    if not real:
        x = torch.complex(torch.rand(K, N, M, generator=g, device="cuda", dtype=rdt) * 2 - 1,
                          torch.rand(K, N, M, generator=g, device="cuda", dtype=rdt) * 2 - 1)
        ref = torch.fft.fft(x, dim=1) if not inverse else torch.fft.ifft(x, dim=1) * N
        nbytes = 2.0 * M * N * K * 2 * fp
        flops = 5.0 * N * math.log2(N) * M * K
    elif not inverse:
        x = torch.rand(K, N, M, generator=g, device="cuda", dtype=rdt) * 2 - 1
        ref = torch.fft.rfft(x, dim=1)
        nbytes = M * K * (N * fp + H * 2.0 * fp)
        flops = 2.5 * N * math.log2(N) * M * K
    else:
        xr = torch.rand(K, N, M, generator=g, device="cuda", dtype=rdt) * 2 - 1
        x = torch.fft.rfft(xr, dim=1)
        ref = xr * N
        nbytes = M * K * (N * fp + H * 2.0 * fp)
        flops = 2.5 * N * math.log2(N) * M * K
    x = x.contiguous()
    if inplace:
        # one buffer in the in-place layout (real rows padded to 2 (N/2 + 1) reals)
        ist, ost = list(cfg.istride[:3]), list(cfg.ostride[:3])
        if not real:
            buf = torch.zeros(ist[2] * K, dtype=cdt, device="cuda")
            inv, outv = buf, buf
        else:
            buf = torch.zeros(2 * H * M * K, dtype=rdt, device="cuda")
            cview = torch.view_as_complex(buf.view(-1, 2))
            inv, outv = (buf, cview) if not inverse else (cview, buf)
        inv.as_strided(x.shape, (ist[2], ist[1], 1)).copy_(x)
        run = lambda: plan.execute(buf)  # noqa: E731
        out_view = lambda: outv.as_strided(ref.shape, (ost[2], ost[1], 1))  # noqa: E731
    else:
        out = torch.empty(ref.shape, dtype=ref.dtype, device="cuda")
        run = lambda: plan.execute(x, out)  # noqa: E731
        out_view = lambda: out  # noqa: E731
    run()
    torch.cuda.synchronize()
    y = out_view()
    err = (torch.linalg.vector_norm((y - ref).reshape(-1).to(torch.complex128 if y.is_complex()
                                                          else torch.float64)) /
           torch.linalg.vector_norm(ref.reshape(-1).to(torch.complex128 if ref.is_complex()
                                                       else torch.float64))).item()
    eps = 1.2e-7 if fp == 4 else 2.3e-16
    if not err <= 8 * eps * max(1.0, math.log2(N)) * 4:
        raise RuntimeError(f"{desc}: result check failed (rel-L2 {err:.3e})")
    run()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        run()
        e.record()
        torch.cuda.synchronize()
        best = min(best, s.elapsed_time(e) * 1e-3)
    return [p, N, int(inplace), M, K, "r" if real else "c", int(inverse), best,
            nbytes / best * 1e-9, flops / best * 1e-9]


def main(argv=None):
    ap = argparse.ArgumentParser(description="double-batched FFT benchmark (CSV)")
    g = ap.add_mutually_exclusive_group()
    g.add_argument("-i", dest="inplace", action="store_true", default=True,
                   help="in-place (default)")
    g.add_argument("-o", dest="inplace", action="store_false", help="out-of-place")
    ap.add_argument("-v", dest="inverse", action="store_true", help="backward transform")
    ap.add_argument("-m", dest="MM", type=int, action="append",
                    help="size of the first mode (repeatable)")
    ap.add_argument("-k", dest="k", type=int, default=0, help="fixed size of the third mode")
    ap.add_argument("-b", dest="b", type=parse_bytes, default=parse_bytes("1G"),
                    help="K from the input tensor's memory footprint (k/M/G suffixes)")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("pd", help="(s|d)(r|c): precision and domain")
    ap.add_argument("NN", type=int, nargs="+", help="sizes of the second mode")
    a = ap.parse_args(argv)
    if len(a.pd) != 2 or a.pd[0] not in "sd" or a.pd[1] not in "rc":
        ap.error("expected (s|d)(r|c)")
    import torch

    import bbfft_amd as bb
    p, d = a.pd
    print(",".join(COLUMNS), flush=True)
    for M in a.MM or [1]:
        for N in a.NN:
            K = k_for(a, M, N, 4 if p == "s" else 8, d == "c")
            row = run_case(bb, torch, p, d, M, N, K, a.inplace, a.inverse, a.reps)
            print(",".join(str(v) for v in row), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
