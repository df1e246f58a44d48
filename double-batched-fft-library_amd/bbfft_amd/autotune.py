"""Autotuner for the gfx950 kernels: picks the radix factorisation, the work-items per
transform, the M x K tile per work-group, the LDS padding and the streaming (non-temporal)
hints for each (precision, transform type, N, M), and writes the table the library compiles in
(double-batched-fft-library_amd/tuning/gfx950.inc).

    python -m bbfft_amd.autotune [-o TABLE] [--rounds 2] DESCRIPTOR...
    python -m bbfft_amd.autotune [-o TABLE] --merge --ragged-serving

Search: coordinate descent from the library's heuristic default, one parameter group at a time
(factorisation + work-items, K tile, M tile, LDS padding, streaming hints, occupancy target,
persistent grid with and without the load-first tile loop), keeping the fastest. Every
candidate's output is compared with the default plan's output on the same input (relative L2
within the precision's tolerance) before it is timed; the default plan itself is covered by the
GPU parity suite against the FP64 oracle. Time = best of 3 runs of `reps` launches (HIP events);
shapes under 1 GB of input plus output cycle through several input/output pairs, so every
launch reads from HBM rather than from the 256 MB last-level cache.
Confirmation (`--confirm ROUNDS`, default 6): the 4 fastest candidates are re-timed in turn,
round after round, over `--confirm-reps` launches each (default 100), and the best median
wins, so that a few per cent of drift between the search's measurements does not decide close
calls and the candidates are ranked at sustained clocks (short bursts can rank LDS-bound
kernels differently).
"""
import argparse
import itertools
import json
import math
import os
import sys
import time

from . import (GFX950_INFO, PREFETCH, BadConfiguration, BbfftError, default_cache,
               kernel_configuration, make_plan, parse_fft_descriptor)

TABLE_HEADER = ("# Tuned kernel parameters for gfx950 (MI355X), written by "
                "`python -m bbfft_amd.autotune`.\n"
                "# precision type N M : factorization Nt Mb Kb lds_pad flags "
                "[waves_per_eu [groups_per_cu]]"
                "   # GB/s default -> tuned, descriptor\n")


def _divisors(n):
    return [d for d in range(1, n + 1) if n % d == 0]


def candidate_factorizations(nfft, fp, per_length=4):
    """Factorisations into 1-4 radices in [2, 32]: the `per_length` most balanced ones for each
    number of stages (largest radix first), fewest stages first."""
    rmax = 32

    def splits(n, parts):
        if parts == 1:
            return [(n,)] if 2 <= n <= rmax else []
        out = []
        for a in _divisors(n):
            if 2 <= a <= rmax:
                out += [(a,) + rest for rest in splits(n // a, parts - 1)]
        return out

    if nfft == 1:
        return [[1]]
    result = []
    for parts in range(1, 5):
        multisets = {tuple(sorted(f, reverse=True)) for f in splits(nfft, parts)}
        ranked = sorted(multisets, key=lambda f: (max(f) - min(f), -f[0]))[:per_length]
        for m in ranked:
            # stage orders: all of them up to 3 stages, largest-first and smallest-first for 4
            orders = sorted(set(itertools.permutations(m)), reverse=True) if parts <= 3 \
                else sorted({m, tuple(reversed(m))}, reverse=True)
            result += [list(o) for o in orders]
    return result


def _pow2_upto(n):
    v, out = 1, []
    while v <= n:
        out.append(v)
        v *= 2
    return out


STORAGE_BITS = 8 | 16 | 32  # half-precision input (f16 / bf16) and the MFMA stage


class Tuner:
    def __init__(self, desc, reps=10, verbose=True, variant=0):
        """`variant`: half-input kernel bits (8 f16, 16 bf16, | 32 MFMA stage) or 0."""
        import torch
        self.torch = torch
        self.desc = desc
        self.cfg = parse_fft_descriptor(desc)
        if self.cfg.dim != 1:
            raise ValueError("autotuning applies to 1-D descriptors")
        self.reps = reps
        self.verbose = verbose
        self.variant = variant & STORAGE_BITS
        self.cache = default_cache()
        M, N, K = self.cfg.shape[:3]
        self.M, self.N, self.K = M, N, K
        self.fp = int(self.cfg.fp)
        t = int(self.cfg.type)
        self.type = t
        real = t != 0
        self.nfft = N // 2 if real and N % 2 == 0 else N
        rdt = torch.float32 if self.fp == 4 else torch.float64
        g = torch.Generator(device="cuda")
        g.manual_seed(N * 7 + M)
        istride, ostride = self.cfg.istride, self.cfg.ostride
        ishape = [M, N if t != 2 else N // 2 + 1, K]
        oshape = [M, N if t != 1 else N // 2 + 1, K]
        n_in = sum((ishape[i] - 1) * istride[i] for i in range(3)) + 1
        n_out = sum((oshape[i] - 1) * ostride[i] for i in range(3)) + 1
        in_elems = n_in if t == 1 else 2 * n_in
        out_elems = n_out if t == 2 else 2 * n_out
        if self.variant:
            # (re, im) half pairs as int32 words; uniform[-1, 1) values
            hdt = torch.float16 if self.variant & 8 else torch.bfloat16
            v = (torch.rand(2 * n_in, generator=g, device="cuda") * 2 - 1).to(hdt)
            self.x = v.view(torch.int32)
            in_bytes = 4 * n_in
        else:
            self.x = torch.rand(in_elems, generator=g, device="cuda", dtype=rdt) * 2 - 1
            in_bytes = in_elems * (4 if self.fp == 4 else 8)
        self.y = torch.empty(out_elems, device="cuda", dtype=rdt)
        self.ref = torch.empty_like(self.y)
        self.bytes = in_bytes + out_elems * (4 if self.fp == 4 else 8)
        # Timing cycles through enough (input, output) pairs that a launch never finds its data
        # in the 256 MB last-level cache (under 1 GB per pair: pairs spanning >= 2 GB).
        npairs = 1 if self.bytes >= ROTATE_BELOW_BYTES else \
            min(64, -(-ROTATE_BYTES // int(self.bytes)))
        self.pairs = [(self.x, self.y)] + [
            (self.x.clone(), torch.empty_like(self.y)) for _ in range(npairs - 1)]
        self.default_ms = None
        self.results = {}

    def log(self, **kw):
        if self.verbose:
            print(json.dumps(kw), flush=True)

    def time_plan(self, plan):
        torch = self.torch
        P = len(self.pairs)
        plan.execute(self.x, self.y)
        plan.execute(self.x, self.y)
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(3):
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record()
            for i in range(self.reps):
                x, y = self.pairs[i % P]
                plan.execute(x, y)
            e.record()
            torch.cuda.synchronize()
            best = min(best, s.elapsed_time(e) / self.reps)
        return best

    def evaluate(self, tuning):
        key = json.dumps(tuning, sort_keys=True)
        if key in self.results:
            return self.results[key]
        torch = self.torch
        try:
            plan = make_plan(self.cfg, cache=self.cache, tuning=tuning)
            self.y.zero_()
            plan.execute(self.x, self.y)
            torch.cuda.synchronize()
            err = (torch.linalg.vector_norm(self.y - self.ref) /
                   torch.linalg.vector_norm(self.ref)).item()
            tol = 8.0 * (1.2e-7 if self.fp == 4 else 2.3e-16) * max(1.0, math.log2(self.N))
            if not err <= tol:
                self.log(desc=self.desc, tuning=tuning, rejected=f"rel err {err:.3g} > {tol:.3g}")
                ms = float("inf")
            else:
                ms = self.time_plan(plan)
        except (BadConfiguration, BbfftError, ValueError) as ex:
            self.log(desc=self.desc, tuning=tuning, skipped=str(ex).splitlines()[0][:160])
            ms = float("inf")
        self.results[key] = ms
        if ms < float("inf"):
            self.log(desc=self.desc, tuning=tuning, ms=round(ms, 5),
                     GBps=round(self.bytes / ms / 1e6, 1))
        return ms

    def full_tuning(self, partial):
        """Complete a partial tuning with the values the library would choose for it (the
        variant bits are part of every tuning's flags)."""
        partial = dict(partial, flags=partial.get("flags", 0) | self.variant)
        kc = kernel_configuration(self.cfg, tuning=partial, device_info=GFX950_INFO)
        return {"factorization": kc["factorization"], "Nt": kc["Nt"], "Mb": kc["Mb"],
                "Kb": kc["Kb"], "lds_pad": kc["lds_pad"], "no_pad": kc["lds_pad"] == 0,
                "flags": partial["flags"], "waves_per_eu": partial.get("waves_per_eu", 0),
                "groups_per_cu": partial.get("groups_per_cu", 0)}

    def factorizations(self):
        if self.variant & 32:  # the MFMA stage is a first radix of 16
            return [[16] + f for f in candidate_factorizations(self.nfft // 16, self.fp)
                    if f != [1]]
        return candidate_factorizations(self.nfft, self.fp)

    def run(self, rounds=2):
        torch = self.torch
        # reference output and time of the heuristic default (explicit, table-independent)
        base = self.full_tuning({"flags": self.variant})
        plan = make_plan(self.cfg, cache=self.cache, tuning=base)
        plan.execute(self.x, self.ref)
        torch.cuda.synchronize()
        self.default_ms = self.time_plan(plan)
        self.results[json.dumps(base, sort_keys=True)] = self.default_ms
        self.log(desc=self.desc, default=base, ms=round(self.default_ms, 5),
                 GBps=round(self.bytes / self.default_ms / 1e6, 1))
        best, best_ms = dict(base), self.default_ms
        kcols = (self.K + 1) // 2 if (self.type != 0 and self.N % 2 == 1) else self.K
        for _ in range(rounds):
            # 1) factorisation and work-items per transform (tile re-derived by the library),
            #    also with a one-column K tile: small groups (down to a single wave, which
            #    needs no barriers) can win even where the default K tile does not
            opts = []
            for f in self.factorizations():
                nt0 = self.nfft // max(f[1:] if self.variant & 32 else f)
                for nt in sorted({nt0, 2 * nt0}):
                    if nt <= 256:
                        o = {"factorization": f, "Nt": nt, "flags": best["flags"],
                             "lds_pad": best["lds_pad"], "no_pad": best["no_pad"]}
                        opts += [o, dict(o, Kb=1)]
            best, best_ms = self._try(best, best_ms, opts, complete=True)
            # 2) K tile
            opts = [dict(best, Kb=kb) for kb in _pow2_upto(min(64, kcols))
                    if best["Mb"] * best["Nt"] * kb <= 1024]
            best, best_ms = self._try(best, best_ms, opts)
            # 3) M tile
            mbs = {self.M} | {max(1, -(-self.M // d)) for d in (2, 4)}
            opts = [dict(best, Mb=mb) for mb in sorted(mbs) if mb * best["Nt"] <= 1024]
            best, best_ms = self._try(best, best_ms, opts)
            # 4) LDS padding
            opts = [dict(best, lds_pad=p, no_pad=p == 0) for p in (0, 8, 16, 32)]
            best, best_ms = self._try(best, best_ms, opts)
            # 5) streaming hints
            opts = [dict(best, flags=fl | (best["flags"] & PREFETCH) | self.variant)
                    for fl in (0, 1, 2, 3)]
            best, best_ms = self._try(best, best_ms, opts)
            # 6) occupancy target of the register allocator
            opts = [dict(best, waves_per_eu=w) for w in (0, 4, 5, 6, 8)]
            best, best_ms = self._try(best, best_ms, opts)
            # 7) persistent grid (work-groups per CU looping over the tiles), with and without
            #    loading the next tile's inputs during the current tile
            pfs = (0, PREFETCH) if self.type != 2 and not self.variant & 32 else (0,)
            opts = [dict(best, groups_per_cu=g, flags=(best["flags"] & ~PREFETCH) | pf)
                    for g in (0, 1, 2, 3, 4, 6, 8, 16) for pf in pfs]
            best, best_ms = self._try(best, best_ms, opts)
        return best, best_ms

    def confirm(self, top=4, rounds=6, reps=None):
        """Re-decide between the `top` fastest distinct candidates of the search by interleaved
        timing in this process: `rounds` rounds, each timing every candidate once in turn over
        `reps` launches (default: the search's), so slow drifts of the device (clocks,
        temperature) hit all of them alike; long runs time the kernels at the sustained clocks
        of a long job. Returns the candidate with the best median and that median (ms)."""
        reps = reps or self.reps
        import statistics
        torch = self.torch
        ranked = sorted((ms, k) for k, ms in self.results.items() if ms < float("inf"))[:top]
        plans = [(json.loads(k), make_plan(self.cfg, cache=self.cache, tuning=json.loads(k)))
                 for _, k in ranked]
        times = [[] for _ in plans]
        P = len(self.pairs)
        for _ in range(rounds):
            for i, (_, plan) in enumerate(plans):
                plan.execute(self.x, self.y)
                s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                s.record()
                for r in range(reps):
                    x, y = self.pairs[r % P]
                    plan.execute(x, y)
                e.record()
                torch.cuda.synchronize()
                times[i].append(s.elapsed_time(e) / reps)
        med = [statistics.median(t) for t in times]
        for (t, _), m in zip(plans, med):
            self.log(desc=self.desc, confirm=t, median_ms=round(m, 5),
                     GBps=round(self.bytes / m / 1e6, 1))
        i = min(range(len(plans)), key=lambda j: med[j])
        return plans[i][0], med[i]

    def _try(self, best, best_ms, options, complete=False):
        tunings = []
        for o in options:
            try:
                t = self.full_tuning(o) if complete else dict(o)
            except (BadConfiguration, BbfftError, ValueError) as ex:
                self.log(desc=self.desc, tuning=o, skipped=str(ex).splitlines()[0][:160])
                continue
            t["flags"] = o.get("flags", 0) | self.variant
            tunings.append(t)
        # compile (hiprtc) the candidates not timed yet in parallel, then time them in turn
        fresh = [t for t in tunings if json.dumps(t, sort_keys=True) not in self.results]
        if len(fresh) > 1:
            from concurrent.futures import ThreadPoolExecutor
            with ThreadPoolExecutor(max_workers=16) as ex:
                list(ex.map(lambda t: _compile_quietly(self, t), fresh))
        for t in tunings:
            ms = self.evaluate(t)
            if ms < best_ms * 0.995:  # keep the incumbent unless clearly faster
                best, best_ms = t, ms
        return best, best_ms


def table_line(cfg, tuning, default_ms, best_ms, nbytes, desc):
    prec = "f32" if int(cfg.fp) == 4 else "f64"
    typ = ["c2c", "r2c", "c2r"][int(cfg.type)]
    flags = tuning["flags"]
    if flags & 8:
        typ += "-f16"
    if flags & 16:
        typ += "-bf16"
    if flags & 32:
        typ += "-mfma"
    fac = "x".join(str(f) for f in tuning["factorization"])
    waves = tuning.get("waves_per_eu", 0)
    groups = tuning.get("groups_per_cu", 0)
    extra = f" {waves} {groups}" if groups else (f" {waves}" if waves else "")
    return (f"{prec} {typ} {cfg.shape[1]} {cfg.shape[0]} : {fac} {tuning['Nt']} {tuning['Mb']} "
            f"{tuning['Kb']} {tuning['lds_pad']} {flags & (3 | PREFETCH)}{extra}   # "
            f"{nbytes / default_ms / 1e6:.0f} -> {nbytes / best_ms / 1e6:.0f} GB/s, {desc}")


def write_table(path, lines):
    with open(path, "w") as f:
        f.write('R"TABLE(\n' + TABLE_HEADER + "\n".join(lines) + '\n)TABLE"\n')


def _key(line):
    """(precision, type, N, M) of a table line, or None for comments / blank lines."""
    parts = line.split("#", 1)[0].split()
    return tuple(parts[:4]) if len(parts) >= 5 and parts[4] == ":" else None


def merge_table(path, new_lines):
    """Entries of the table at `path` with those of `new_lines` replacing equal keys."""
    old = []
    if os.path.exists(path):
        with open(path) as f:
            text = f.read()
        body = text.split('R"TABLE(', 1)[-1].split(')TABLE"', 1)[0]
        old = [ln for ln in body.splitlines() if _key(ln)]
    keys = {_key(ln) for ln in new_lines}
    return [ln for ln in old if _key(ln) not in keys] + list(new_lines)


ROTATE_BELOW_BYTES = 1 << 30  # (input + output) bytes below which timing rotates buffers
ROTATE_BYTES = 2 << 30  # bytes the rotated pairs span together
RAGGED_GROUP = 256  # work-items of the persistent ragged kernel's work-group
RAGGED_WAVES = 5  # its occupancy target (waves per SIMD), shared by all variants


def ragged_candidates(nfft, M, rmax=16, group=RAGGED_GROUP):
    """(factorisation, Nt) pairs for a ragged variant: up to three stages, radices <= rmax,
    one transform's work-items times M within the common work-group."""
    out = []
    for f in candidate_factorizations(nfft, 4, per_length=3):
        if len(f) > 3 or max(f) > rmax:
            continue
        nt = nfft // max(f)
        if M * nt <= group:
            out.append((f, nt))
    return out


def _ragged_tuning(f, nt, M, pad, flags):
    return {"factorization": list(f), "Nt": nt, "Mb": M, "Kb": max(1, RAGGED_GROUP // (M * nt)),
            "lds_pad": pad, "no_pad": pad == 0, "flags": flags,
            "waves_per_eu": RAGGED_WAVES, "groups_per_cu": 0}


def tune_ragged_serving(reps=20, confirm_rounds=4, confirm_reps=50, elements=1 << 25,
                        workers=16, verbose=True):
    """Tune the variants (N, M) of the seeded ragged serving workload
    (bbfft_amd.ragged.synthetic_serving_jobs) under the persistent kernel's constraints: each
    candidate (factorisation with radices <= 16, work-items per transform, LDS padding, access
    hints) runs as a plain kernel with the ragged kernel's work-group size and occupancy target
    on `elements` complex elements, so register pressure that would spill in the shared kernel
    shows as a slow candidate. Direction does not change the cost, so forward and backward share
    an entry. Candidates are compiled in parallel first. Returns the table lines
    ("f32 c2c-ragged N M : ...") ordered by the variants' share of the workload."""
    from concurrent.futures import ThreadPoolExecutor
    from . import ragged
    jobs = ragged.synthetic_serving_jobs()
    var = ragged.job_variants(jobs)
    weight, current = {}, {}
    for j, v in zip(jobs, var):
        weight[(j.N, j.M)] = weight.get((j.N, j.M), 0) + j.elements
        if (j.N, j.M) not in current:
            current[(j.N, j.M)] = ragged.variant_configuration(jobs, v)
    total = sum(weight.values())
    lines = []
    for (N, M), w in sorted(weight.items(), key=lambda kv: -kv[1]):
        t0 = time.time()
        K = max(64, elements // (M * N))
        desc = f"scfo{M}.{N}*{K}"
        tuner = Tuner(desc, reps=reps, verbose=verbose)
        cur = current[(N, M)]
        base = _ragged_tuning(cur["factorization"], cur["Nt"], cur["Mb"], cur["lds_pad"], 0)
        base["Kb"] = cur["Kb"]
        # reference output: the library's plain plan for the shape
        plan = make_plan(tuner.cfg, cache=tuner.cache)
        plan.execute(tuner.x, tuner.ref)
        tuner.torch.cuda.synchronize()
        cands = [_ragged_tuning(f, nt, M, 16, 0) for f, nt in ragged_candidates(tuner.nfft, M)]
        with ThreadPoolExecutor(max_workers=workers) as ex:  # compile (hiprtc) in parallel
            list(ex.map(lambda t: _compile_quietly(tuner, t), [base] + cands))
        tuner.default_ms = tuner.evaluate(base)
        for t in cands:
            tuner.evaluate(t)
        # padding and access hints for the two fastest factorisations
        top = sorted((ms, k) for k, ms in tuner.results.items() if ms < float("inf"))[:2]
        more = []
        for _, k in top:
            t = json.loads(k)
            for pad in (0, 16):
                for fl in (0, 1, 2, 3):
                    more.append(_ragged_tuning(t["factorization"], t["Nt"], M, pad, fl))
        with ThreadPoolExecutor(max_workers=workers) as ex:
            list(ex.map(lambda t: _compile_quietly(tuner, t), more))
        for t in more:
            tuner.evaluate(t)
        best, best_ms = tuner.confirm(rounds=confirm_rounds, reps=confirm_reps)
        fac = "x".join(str(f) for f in best["factorization"])
        line = (f"f32 c2c-ragged {N} {M} : {fac} {best['Nt']} {best['Mb']} {best['Kb']} "
                f"{best['lds_pad']} {best['flags'] & 3} {RAGGED_WAVES}   # "
                f"{tuner.bytes / tuner.default_ms / 1e6:.0f} -> {tuner.bytes / best_ms / 1e6:.0f}"
                f" GB/s, {desc} as ragged variant ({100.0 * w / total:.1f} % of the serving "
                f"workload)")
        lines.append(line)
        print(json.dumps({"variant": [N, M], "best": best, "ms": round(best_ms, 5),
                          "current_ms": round(tuner.default_ms, 5),
                          "seconds": round(time.time() - t0, 1), "line": line}), flush=True)
        del tuner
    return lines


def _compile_quietly(tuner, tuning):
    """Build (and cache) the plan of a candidate; failures are left to evaluate()."""
    try:
        make_plan(tuner.cfg, cache=tuner.cache, tuning=tuning)
    except (BadConfiguration, BbfftError, ValueError):
        pass


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("-o", "--output", default=None, help="table file to write (.inc)")
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--merge", action="store_true",
                    help="keep the other entries of an existing output table")
    ap.add_argument("--storage", choices=["fp32", "f16", "bf16"], default="fp32",
                    help="input storage of c2c fp32 descriptors (half-input kernels)")
    ap.add_argument("--mfma", action="store_true", help="with --storage f16/bf16: MFMA stage")
    ap.add_argument("--confirm", type=int, default=6, metavar="ROUNDS",
                    help="re-time the 4 fastest candidates interleaved over ROUNDS rounds and "
                         "keep the best median (0: keep the search's best)")
    ap.add_argument("--confirm-reps", type=int, default=100, metavar="N",
                    help="launches per candidate and round of the confirmation; long runs rank "
                         "the candidates at the sustained clocks of a long job (0: --reps)")
    ap.add_argument("--ragged-serving", action="store_true",
                    help="tune the variants of the ragged serving workload instead of "
                         "descriptors (table entries c2c-ragged)")
    ap.add_argument("descriptors", nargs="*")
    a = ap.parse_args(argv)
    if a.ragged_serving:
        lines = tune_ragged_serving(reps=a.reps, confirm_rounds=a.confirm,
                                    confirm_reps=a.confirm_reps or a.reps)
        if a.output:
            write_table(a.output, merge_table(a.output, lines) if a.merge else lines)
        return 0
    if not a.descriptors:
        ap.error("descriptors are required (or --ragged-serving)")
    lines = []
    for desc in a.descriptors:
        t0 = time.time()
        variant = {"fp32": 0, "f16": 8, "bf16": 16}[a.storage]
        if variant and a.mfma:
            variant |= 32
        tuner = Tuner(desc, reps=a.reps, variant=variant)
        best, best_ms = tuner.run(a.rounds)
        if a.confirm > 0:
            best, best_ms = tuner.confirm(rounds=a.confirm, reps=a.confirm_reps or None)
        line = table_line(tuner.cfg, best, tuner.default_ms, best_ms, tuner.bytes, desc)
        lines.append(line)
        print(json.dumps({"desc": desc, "best": best, "ms": round(best_ms, 5),
                          "default_ms": round(tuner.default_ms, 5),
                          "speedup": round(tuner.default_ms / best_ms, 4),
                          "seconds": round(time.time() - t0, 1), "line": line}), flush=True)
        if a.output:
            write_table(a.output, merge_table(a.output, lines) if a.merge else lines)
        del tuner
    return 0


if __name__ == "__main__":
    sys.exit(main())
