"""Ragged batches: many c2c jobs of different N, M, K and direction in one persistent launch
(binding of the bbfft_ragged_* C entry points; no reference counterpart).

    jobs = ragged.pack([ragged.Job(M=2, N=60, K=1000, direction=-1), ...])
    plan = ragged.make_ragged_plan(jobs, precision=4, graph=True)
    plan.execute(x, y)        # x, y: device buffers of ragged.total_elements(jobs) complex

A job is an M x N x K complex tensor with packed column-major strides {1, M, M*N} that starts
`in_offset` / `out_offset` complex elements after the buffers given to execute.
"""
import ctypes
import math
from dataclasses import dataclass, replace
from typing import List, Sequence

import numpy as np

from . import GFX950_INFO, Cache, Plan, _capi, _check, _current_stream

# Lengths of the synthetic serving workload (mixed powers of two, smooth and odd lengths).
SERVING_LENGTHS = (12, 60, 100, 128, 243, 384, 500, 512)


@dataclass(frozen=True)
class Job:
    M: int
    N: int
    K: int
    direction: int = -1
    in_offset: int = 0
    out_offset: int = 0

    @property
    def elements(self) -> int:
        return self.M * self.N * self.K

    def descriptor(self, precision=4) -> str:
        """Descriptor of the job as a plain out-of-place plan (e.g. "scfo2.60*1000")."""
        p = "s" if precision == 4 else "d"
        d = "f" if self.direction < 0 else "b"
        return f"{p}c{d}o{self.M}.{self.N}*{self.K}"


def _c_jobs(jobs: Sequence[Job]):
    arr = (_capi.CRaggedJob * max(1, len(jobs)))()
    for i, j in enumerate(jobs):
        arr[i] = _capi.CRaggedJob(j.M, j.N, j.K, int(j.direction), j.in_offset, j.out_offset)
    return arr


def pack(jobs: Sequence[Job]) -> List[Job]:
    """Consecutive offsets in job order (in_offset == out_offset: in-place capable layout)."""
    out, off = [], 0
    for j in jobs:
        out.append(replace(j, in_offset=off, out_offset=off))
        off += j.elements
    return out


def total_elements(jobs: Sequence[Job]) -> int:
    """Complex elements a buffer needs to hold every job's input (or output)."""
    return max((max(j.in_offset, j.out_offset) + j.elements for j in jobs), default=0)


def synthetic_serving_jobs(n_jobs=4096, seed=2024, lengths=SERVING_LENGTHS, k_min=64,
                           k_max=65536, m_max=4) -> List[Job]:
    """Seeded mixed workload: N uniform over `lengths`, K log-uniform in [k_min, k_max], M
    uniform in 1..m_max, directions alternating forward / backward; packed offsets."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ns = rng.integers(0, len(lengths), n_jobs)
    ks = np.exp(rng.uniform(math.log(k_min), math.log(k_max), n_jobs))
    ms = rng.integers(1, m_max + 1, n_jobs)
    jobs = [Job(M=int(ms[i]), N=int(lengths[ns[i]]),
                K=int(min(k_max, max(k_min, round(float(ks[i]))))),
                direction=-1 if i % 2 == 0 else 1) for i in range(n_jobs)]
    return pack(jobs)


def kernel_configuration(jobs, precision=4, device_info=GFX950_INFO) -> dict:
    L = _capi.load()
    info = _capi.CRaggedInfo()
    name = ctypes.create_string_buffer(256)
    _check(L.bbfft_ragged_kernel_configuration(precision, _c_jobs(jobs), len(jobs),
                                               device_info.encode(), ctypes.byref(info), name,
                                               256))
    return {"group_size": info.group_size, "lds_bytes": info.lds_bytes,
            "num_variants": info.num_variants, "total_tiles": info.total_tiles,
            "name": name.value.decode()}


def variant_configuration(jobs, variant, precision=4, device_info=GFX950_INFO) -> dict:
    L = _capi.load()
    kc = _capi.CKernelConfig()
    _check(L.bbfft_ragged_variant_configuration(precision, _c_jobs(jobs), len(jobs),
                                                device_info.encode(), variant,
                                                ctypes.byref(kc)))
    return {"M": kc.M, "Mb": kc.Mb, "N": kc.N, "Nt": kc.Nt, "Kb": kc.Kb,
            "lds_pad": kc.lds_pad, "lds_bytes": kc.lds_bytes,
            "work_group_size": kc.work_group_size,
            "factorization": [kc.factorization[i] for i in range(kc.num_factors)]}


def job_variants(jobs, precision=4, device_info=GFX950_INFO) -> List[int]:
    L = _capi.load()
    out = (ctypes.c_uint * max(1, len(jobs)))()
    _check(L.bbfft_ragged_job_variants(precision, _c_jobs(jobs), len(jobs),
                                       device_info.encode(), out))
    return list(out)[:len(jobs)]


def job_table(jobs, precision=4, device_info=GFX950_INFO, grid=1) -> np.ndarray:
    """Device job table for a launch of `grid` work-groups (see bbfft_ragged_job_table)."""
    L = _capi.load()
    cj = _c_jobs(jobs)
    n = ctypes.c_size_t()
    _check(L.bbfft_ragged_job_table(precision, cj, len(jobs), device_info.encode(), grid, None,
                                    0, ctypes.byref(n)))
    t = np.zeros(n.value, dtype=np.uint64)
    _check(L.bbfft_ragged_job_table(precision, cj, len(jobs), device_info.encode(), grid,
                                    t.ctypes.data_as(ctypes.POINTER(ctypes.c_ulonglong)),
                                    n.value, ctypes.byref(n)))
    return t


def generate_source(jobs, precision=4, device_info=GFX950_INFO) -> str:
    L = _capi.load()
    cj = _c_jobs(jobs)
    n = ctypes.c_size_t()
    _check(L.bbfft_generate_ragged_source(precision, cj, len(jobs), device_info.encode(), None,
                                          0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value)
    _check(L.bbfft_generate_ragged_source(precision, cj, len(jobs), device_info.encode(), buf,
                                          n.value, ctypes.byref(n)))
    return buf.value.decode()


def aot_generate(jobs, out_path, precision=4, device_info=GFX950_INFO, arch="gfx950"):
    """Compile the ragged kernel for these jobs' variants into an archive (no GPU needed)."""
    _check(_capi.load().bbfft_aot_generate_ragged(precision, _c_jobs(jobs), len(jobs),
                                                  device_info.encode(), arch.encode(),
                                                  out_path.encode()))


def make_ragged_plan(jobs, precision=4, stream=None, cache: Cache = None, graph=True) -> Plan:
    """Plan running every job in one persistent launch on the current HIP device."""
    if stream is None:
        stream = _current_stream()
    h = ctypes.c_void_p()
    flags = _capi.RAGGED_GRAPH if graph else 0
    ch = cache.handle if cache is not None else None
    _check(_capi.load().bbfft_ragged_plan_create(precision, _c_jobs(jobs), len(jobs), flags,
                                                 stream, ch, ctypes.byref(h)))
    return Plan(h, None, stream, io_bytes=io_bytes(jobs, precision))


def io_bytes(jobs: Sequence[Job], precision=4):
    """(input, output) bytes a ragged launch addresses from the buffers' start: every job
    reads [in_offset, in_offset + M N K) and writes [out_offset, out_offset + M N K) complex
    elements."""
    eb = 2 * precision
    in_end = max((j.in_offset + j.elements for j in jobs), default=0)
    out_end = max((j.out_offset + j.elements for j in jobs), default=0)
    return in_end * eb, out_end * eb
