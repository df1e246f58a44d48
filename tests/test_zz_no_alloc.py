"""GPU tests of bbfft_plan_execute_static_scratch, the allocation-free variant of
bbfft_plan_execute (no counterpart in the reference): its output must equal bbfft_plan_execute's
bit for bit, call after call, since both launch the same kernels and only the scratch buffer
between the passes differs (the plan's own allocation vs the library's per-GPU static array)."""
import ctypes

import numpy as np
import pytest

import bbfft_amd as bb
import gpu_cases
from bbfft_amd import _capi
from gpu_util import have_gpu, io_shapes, extent

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not have_gpu(), reason="needs a HIP GPU")]

if have_gpu():
    import torch


def _buffers(cfg, seed):
    """Seeded input tensor of cfg's input layout and a NaN-filled output tensor."""
    ishape, oshape = io_shapes(cfg)
    n_in = extent(cfg.dim, ishape, cfg.istride)
    n_out = extent(cfg.dim, oshape, cfg.ostride)
    rng = np.random.default_rng(seed)
    f32 = int(cfg.fp) == 4
    t = int(cfg.type)
    if t == 1:
        x = rng.uniform(-1, 1, n_in)
    else:
        x = rng.uniform(-1, 1, n_in) + 1j * rng.uniform(-1, 1, n_in)
    x = x.astype((np.float32 if t == 1 else np.complex64) if f32 else
                 (np.float64 if t == 1 else np.complex128))
    xd = torch.from_numpy(x).cuda()
    odt = (torch.float32 if t == 2 else torch.complex64) if f32 else \
        (torch.float64 if t == 2 else torch.complex128)
    yd = torch.full((n_out,), float("nan"), dtype=odt, device="cuda") if t == 2 else \
        torch.full((n_out,), complex(float("nan"), float("nan")), dtype=odt, device="cuda")
    return xd, yd


def _scratch_bytes(cfg):
    """Scratch a 2-D/3-D plan needs: its input size when that exceeds its output size."""
    rb = int(cfg.fp)
    t = int(cfg.type)
    D = cfg.dim
    ib = rb if t == 1 else 2 * rb
    ob = rb if t == 2 else 2 * rb
    isize = cfg.istride[D + 1] * cfg.shape[D + 1] * ib
    osize = cfg.ostride[D + 1] * cfg.shape[D + 1] * ob
    return isize if D > 1 and isize > osize else 0


# Two calls in a row, the second on a smaller input: out-of-place c2r plans (scratch between
# the passes, fused and unfused), then plans without scratch (1-D, r2c, in-place c2r).
_D = gpu_cases.static_scratch()
SEQUENCES = [(_D[i], _D[i + 1]) for i in range(0, 10, 2)]


@pytest.mark.parametrize("seq", SEQUENCES, ids=["+".join(s) for s in SEQUENCES])
def test_static_scratch_matches_execute(seq):
    cache = bb.default_cache()
    for i, desc in enumerate(seq):
        print(f"static scratch: {desc}", flush=True)
        cfg = bb.parse_fft_descriptor(desc)
        plan = bb.make_plan(cfg, cache=cache)
        if desc[3] == "i":
            ishape, oshape = io_shapes(cfg)
            n = max(extent(cfg.dim, ishape, cfg.istride) * (1 if int(cfg.type) == 1 else 2),
                    extent(cfg.dim, oshape, cfg.ostride) * (1 if int(cfg.type) == 2 else 2))
            rb = torch.float32 if int(cfg.fp) == 4 else torch.float64
            raw = torch.from_numpy(np.random.default_rng(3 + i).uniform(-1, 1, n)).to(rb).cuda()
            a, b = raw.clone(), raw.clone()
            plan.execute(a)
            plan.execute(b, static_scratch=True)
            torch.cuda.synchronize()
            assert torch.equal(a, b), desc
            continue
        xd, ref = _buffers(cfg, 7 + i)
        plan.execute(xd, ref)
        torch.cuda.synchronize()
        for rep in range(2):
            got = torch.full_like(ref, float("nan"))
            plan.execute(xd, got, static_scratch=True)
            torch.cuda.synchronize()
            assert torch.equal(torch.view_as_real(got) if got.is_complex() else got,
                               torch.view_as_real(ref) if ref.is_complex() else ref), \
                f"{desc} (call {rep})"
        assert _scratch_bytes(cfg) <= 64 << 20


def test_static_scratch_rejects_oversized_plan():
    """A plan needing more scratch than the 64 MiB array is rejected before any launch."""
    desc = "srbo512x512*64"  # 257 * 512 * 64 complex64 = 67 371 008 bytes of scratch
    cfg = bb.parse_fft_descriptor(desc)
    assert _scratch_bytes(cfg) > 64 << 20
    plan = bb.make_plan(cfg, cache=bb.default_cache())
    xd, yd = _buffers(cfg, 1)
    L = _capi.load()
    rc = L.bbfft_plan_execute_static_scratch(plan._h, xd.data_ptr(), yd.data_ptr(), None, 0,
                                             None)
    torch.cuda.synchronize()
    assert rc == _capi.ERROR_BAD_CONFIGURATION
    msg = L.bbfft_last_error().decode()
    assert "static scratch" in msg and "67371008" in msg, msg
    assert torch.isnan(torch.view_as_real(yd) if yd.is_complex() else yd).all()
    # the allocating entry point still runs the same plan
    plan.execute(xd, yd)
    torch.cuda.synchronize()
    assert torch.isfinite(yd).all()


def test_static_scratch_events():
    """Wait and signal events work as in bbfft_plan_execute."""
    cfg = bb.parse_fft_descriptor("srbo32x16*10")
    plan = bb.make_plan(cfg, cache=bb.default_cache())
    xd, ref = _buffers(cfg, 5)
    plan.execute(xd, ref)
    ev_wait = torch.cuda.Event()
    ev_wait.record()
    ev_signal = torch.cuda.Event()
    ev_signal.record()  # creates the underlying hipEvent_t
    got = torch.full_like(ref, float("nan"))
    plan.execute(xd, got, signal=ev_signal.cuda_event, wait=[ev_wait.cuda_event],
                 static_scratch=True)
    ev_signal.synchronize()
    assert torch.equal(got, ref)
    assert ctypes.c_void_p(ev_signal.cuda_event).value
