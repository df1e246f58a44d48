"""Case lists of the GPU parity tests (pure data).

They restate the shapes of the reference's device tests (test/c2c.cpp, test/r2c.cpp) and the
sizes of the benchmark configurations. __graft_entry__.build() compiles the kernels of all
these descriptors ahead of time so the GPU run does not JIT them.
"""


def desc(prec, domain, dirn, place, M, Ns, K, istride=None, ostride=None):
    """Descriptor string with explicit M and K, e.g. desc('s','c','f','o',8,[256],32768)."""
    s = f"{prec}{domain}{dirn}{place}{M}." + "x".join(str(n) for n in Ns) + f"*{K}"
    if istride is not None:
        s += "i" + ",".join(str(v) for v in istride)
    if ostride is not None:
        s += "o" + ",".join(str(v) for v in ostride)
    return s


PRECS = ["s", "d"]

# test/c2c.cpp:56-63
C2C_KA_M = [1, 2, 3, 16, 17, 64, 256, 1024]
C2C_KA_N = [2, 3, 5, 7, 11, 13, 4, 8, 16, 32, 128, 256, 512, 27, 63, 105, 363]
C2C_KA_K = [1, 32]


def c2c_known_answer():
    return [desc(p, "c", "f", "i", M, [N], K)
            for p in PRECS for K in C2C_KA_K for N in C2C_KA_N for M in C2C_KA_M]


def c2c_nonpacked():
    # test/c2c.cpp:69-83: stride {1, M+1, (M+1)(N+1)} in and out, in place
    out = []
    for p in PRECS:
        for M in [1, 32]:
            for N in [6, 17, 102]:
                s = [1, M + 1, (M + 1) * (N + 1)]
                out.append(desc(p, "c", "f", "i", M, [N], 33, s, s))
    return out


def c2c_identity():
    # test/c2c.cpp:85-127
    return [desc(p, "c", d, "i", M, [N], 16)
            for p in PRECS for d in "fb" for M in [1, 32] for N in [5, 63, 92]]


# Benchmark sweep (M = 8, K = 32768) and its non-powers of two
BENCH_N = [16, 32, 64, 128, 256, 512, 60, 105, 243, 384]


def c2c_random():
    """Random inputs vs the FP64 oracle, both directions, small K."""
    out = []
    for p in PRECS:
        for d in "fb":
            for N in BENCH_N + [12, 100, 500, 17, 363]:
                out.append(desc(p, "c", d, "o", 8, [N], 5))
            for M in [1, 3]:
                for N in [64, 243]:
                    out.append(desc(p, "c", d, "o", M, [N], 7))
    return out


def c1_full_k():
    """The CPU configuration's shape, c2c M=1 x N=64 x K=16384 (full K), both precisions and
    directions, against the FP64 oracle."""
    return [desc(p, "c", d, "o", 1, [64], 16384) for p in PRECS for d in "fb"]


# lengths with a prime factor above the in-register radix limit (direct DFT stage)
DIRECT_N = [67, 134, 254, 262, 509]


def direct_stage():
    """c2c both directions, r2c and c2r, both precisions, at lengths needing a direct stage."""
    out = []
    for p in PRECS:
        for N in DIRECT_N:
            for M in [1, 8]:
                out += [desc(p, "c", d, "o", M, [N], 33) for d in "fb"]
            out += [desc(p, "r", "f", "o", 3, [N], 33), desc(p, "r", "b", "o", 3, [N], 33)]
        out += [desc(p, "c", "f", "o", 2, [134, 6], 3), desc(p, "r", "f", "o", 2, [254, 5], 3),
                desc(p, "r", "b", "o", 2, [262, 5], 3)]
    return out


def direct_full_size():
    return [desc(p, "c", d, "o", 8, [509], 32768) for p in PRECS for d in "fb"]


def c2c_full_size():
    return [desc(p, "c", d, "o", 8, [N], 32768) for p in PRECS for d in "fb" for N in BENCH_N]


def r2c_1d_oop():
    # test/r2c.cpp:171-182
    return [desc(p, "r", "f", "o", M, [N], K) for p in PRECS for K in [1, 33] for M in [1, 3, 32]
            for N in [2, 4, 5, 8, 27, 16, 32, 128, 105, 256, 512, 102, 220, 10, 26]]


def r2c_1d_inplace():
    # test/r2c.cpp:184-194
    return [desc(p, "r", "f", "i", M, [N], 65) for p in PRECS for M in [1, 3]
            for N in [4, 12, 13, 110]]


def r2c_nd():
    # test/r2c.cpp:196-244
    out = []
    for p in PRECS:
        for K in [1, 33]:
            for M in [1, 3]:
                out += [desc(p, "r", "f", "i", M, N, K) for N in [[4, 8], [8, 5]]]
                out += [desc(p, "r", "f", "i", M, N, K) for N in [[4, 8, 2], [8, 256, 5]]]
        for K in [1, 65]:
            for M in [1, 7]:
                out += [desc(p, "r", "f", "o", M, N, K) for N in [[5, 4], [10, 3]]]
                out += [desc(p, "r", "f", "o", M, N, K) for N in [[5, 4, 6], [10, 286, 3]]]
    return out


def c2r_1d_oop():
    # test/r2c.cpp:300-308
    return [desc(p, "r", "b", "o", M, [N], 33) for p in PRECS for M in [1, 5, 32]
            for N in [11, 16, 25, 96, 256, 300, 315]]


def c2r_polluted():
    # test/r2c.cpp:310-324
    return [desc(p, "r", "b", "i", 1, [N], 33) for p in PRECS for N in [5, 16, 21, 48, 512]]


def c2r_nd():
    # test/r2c.cpp:326-374
    out = []
    for p in PRECS:
        for K in [1, 54]:
            for M in [1, 3]:
                out += [desc(p, "r", "b", "i", M, N, K) for N in [[4, 2], [96, 96]]]
                out += [desc(p, "r", "b", "i", M, N, K) for N in [[4, 8, 2], [96, 96, 80]]]
        for K in [1, 65]:
            for M in [1, 7]:
                out += [desc(p, "r", "b", "o", M, N, K) for N in [[5, 4], [10, 3]]]
                out += [desc(p, "r", "b", "o", M, N, K) for N in [[5, 4, 6], [10, 286, 3]]]
    return out


def r2c_identity():
    # test/r2c.cpp:458-497: (M, N, K, inplace)
    out = []
    for p in PRECS:
        for place in "io":
            for M in [1, 8]:
                for N in [[8], [64], [212]]:
                    out.append((p, place, M, N, 16))
        for M in [1, 7]:
            for K in [1, 11]:
                for N in [[8, 8, 8], [96, 96, 80]]:
                    out.append((p, "i", M, N, K))
    return out


def identity_descriptors(case):
    p, place, M, N, K = case
    return [desc(p, "r", "f", place, M, N, K), desc(p, "r", "b", place, M, N, K)]


def c3_descriptors():
    """3-D r2c/c2r of 32 x 128^3 (out-of-place), and a 2-batch version for oracle checks."""
    return [desc("s", "r", d, "o", 1, [128, 128, 128], K) for d in "fb" for K in [32, 2]]


def static_scratch():
    """Plans run through the allocation-free execute entry point (pairs: larger, smaller)."""
    return ["srbo128x96*40", "srbo24x20x12*6", "drbo64x48x10*5", "drbo16x12*9",
            "srbo128x128x128*2", "srbo128x128*8", "scfo8.256*4096", "scbo1.60*77",
            "srfo64x32*20", "srbi32x16x8*3", "srbo512x512*64", "srbo32x16*10"]


def all_descriptors():
    out = (c2c_known_answer() + c2c_nonpacked() + c2c_identity() + c2c_random() +
           c1_full_k() + direct_stage() + direct_full_size() + c2c_full_size() +
           r2c_1d_oop() + r2c_1d_inplace() + r2c_nd() +
           c2r_1d_oop() + c2r_polluted() + c2r_nd() + c3_descriptors() + static_scratch())
    for c in r2c_identity():
        out += identity_descriptors(c)
    seen, uniq = set(), []
    for d in out:
        if d not in seen:
            seen.add(d)
            uniq.append(d)
    return uniq
