"""Transpose-FFT-transpose against one M-batched plan (the non-unit-stride comparison).

    python -m bbfft_amd.tft [-s|-d] [-k K] [--reps R] [M.N ...]

Reference: examples/transpose_fft_transpose/tft.cpp:120-199. For each M x N (default the
reference's double-precision set 16.16, 1120.32, 128.512, 70.16) an M x N x K complex tensor
(K from a 512 MB footprint unless -k is given) is transformed along its second mode twice:

* transpose to N x M x K, one in-place unit-stride plan {1, N, M*K}, transpose back;
* one in-place M-batched plan {M, N, K} on the original layout.

Both results are compared, then each step is timed on its own (minimum of `--reps` runs,
HIP events on the current stream) and reported as GB/s of 2 x the tensor's bytes, like the
reference. The transposes are torch copies of permuted views: they are the comparison's
reference point, not part of the library.
"""
import argparse
import sys

DEFAULT_SHAPES = ["16.16", "1120.32", "128.512", "70.16"]
DEFAULT_TENSOR_BYTES = 512 * 1000 * 1000


def _time(torch, fn, reps):
    best = float("inf")
    for _ in range(reps):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        torch.cuda.synchronize()
        best = min(best, s.elapsed_time(e) * 1e-3)
    return best


def run_case(bb, torch, M, N, K, fp, reps):
    """Returns a dict of the five timings (s), their GB/s and the relative L2 difference."""
    cdt = torch.complex64 if fp == bb.precision.f32 else torch.complex128
    rdt = torch.float32 if fp == bb.precision.f32 else torch.float64
    if K == 0:
        K = max(1, DEFAULT_TENSOR_BYTES // (2 * rdt.itemsize * M * N))
    # bbfft's column-major M x N x K tensor is torch's row-major (K, N, M)
    k = torch.arange(K, device="cuda", dtype=rdt).view(K, 1, 1) / K
    n = torch.arange(N, device="cuda", dtype=rdt).view(1, N, 1) / N
    m = torch.arange(M, device="cuda", dtype=rdt).view(1, 1, M) / M
    x0 = torch.complex(m + n + k, torch.zeros((), device="cuda", dtype=rdt)).contiguous()
    x = torch.empty_like(x0)
    xt = torch.empty((K, M, N), device="cuda", dtype=cdt)
    X = torch.empty_like(x0)

    plan_tft = bb.make_plan(bb.configuration(1, [1, N, M * K], fp, bb.direction.forward,
                                             bb.transform_type.c2c))
    plan_nu = bb.make_plan(bb.configuration(1, [M, N, K], fp, bb.direction.forward,
                                            bb.transform_type.c2c))

    def transpose1():
        xt.copy_(x.transpose(1, 2))

    def transpose2():
        X.copy_(xt.transpose(1, 2))

    def unit_fft():
        plan_tft.execute(xt)

    def tft():
        transpose1()
        unit_fft()
        transpose2()

    def non_unit():
        plan_nu.execute(x)

    x.copy_(x0)
    tft()
    non_unit()
    torch.cuda.synchronize()
    wide = torch.complex128
    err = (torch.linalg.vector_norm((X.to(wide) - x.to(wide)).reshape(-1)) /
           torch.linalg.vector_norm(x.to(wide).reshape(-1))).item()

    x.copy_(x0)
    t = {"transpose1": _time(torch, transpose1, reps), "unit_fft": _time(torch, unit_fft, reps),
         "transpose2": _time(torch, transpose2, reps), "tft": _time(torch, tft, reps)}
    x.copy_(x0)
    t["non_unit"] = _time(torch, non_unit, reps)
    nbytes = 2.0 * x.element_size() * x.numel()
    return {"M": M, "N": N, "K": K, "rel_l2": err, "seconds": t,
            "GBps": {key: nbytes / v * 1e-9 for key, v in t.items()},
            "speedup": t["tft"] / t["non_unit"]}


def main(argv=None):
    ap = argparse.ArgumentParser(description="transpose-FFT-transpose vs M-batched plan")
    g = ap.add_mutually_exclusive_group()
    g.add_argument("-s", dest="double", action="store_false", help="single precision")
    g.add_argument("-d", dest="double", action="store_true", default=True,
                   help="double precision (default, like the reference example)")
    ap.add_argument("-k", dest="k", type=int, default=0, help="fixed K (default: 512 MB tensor)")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("shapes", nargs="*", default=DEFAULT_SHAPES, help="M.N pairs")
    a = ap.parse_args(argv)
    import torch

    import bbfft_amd as bb
    fp = bb.precision.f64 if a.double else bb.precision.f32
    tol = 1e-13 if a.double else 1e-5
    status = 0
    for s in a.shapes:
        M, N = (int(v) for v in s.split("."))
        r = run_case(bb, torch, M, N, a.k, fp, a.reps)
        print(f"{M} x {N} x {r['K']}")
        names = [("transpose1", "Transpose 1"), ("unit_fft", "Unit-stride FFT"),
                 ("transpose2", "Transpose 2"), ("tft", "Transpose-FFT-transpose:"),
                 ("non_unit", "Non-unit stride FFT:")]
        for key, label in names:
            print(f"{label} {r['seconds'][key]:.6g} s, {r['GBps'][key]:.1f} GB / s")
        print(f"Speed-up: {r['speedup']:.3f}x (rel-L2 difference {r['rel_l2']:.2e})",
              flush=True)
        if not r["rel_l2"] <= tol:
            print(f"Error: results differ (rel-L2 {r['rel_l2']:.3e} > {tol:g})")
            status = 1
    return status


if __name__ == "__main__":
    sys.exit(main())
