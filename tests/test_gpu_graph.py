"""Plans inside HIP graphs: a plan's launches (1-D, 2-D/3-D with a scratch buffer, fused
passes) captured with torch.cuda.graph on the plan's stream and replayed several times, with
eager work and synchronisations between the replays; every replay must equal an eager
execution on the same input. Plans allocate nothing and clear nothing at execute time, so
their launches are capturable (reference: plan_unmanaged_event chaining, plan.hpp:144-184)."""
import pytest

import bbfft_amd as bb
from gpu_util import have_gpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not have_gpu(), reason="needs a HIP GPU")]

if have_gpu():
    import torch


def _buffers(cfg, seed):
    D = cfg.dim
    Ns = list(cfg.shape[1:D + 1])
    M, B = cfg.shape[0], cfg.shape[D + 1]
    t = int(cfg.type)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    shape = [B] + Ns[::-1] + [M]
    half = list(shape)
    half[-2] = Ns[0] // 2 + 1
    if t == 0:
        x = torch.randn(shape + [2], generator=g, device="cuda")
        x = torch.view_as_complex(x)
        y = torch.empty_like(x)
    elif t == 1:
        x = torch.randn(shape, generator=g, device="cuda")
        y = torch.empty(half, dtype=torch.complex64, device="cuda")
    else:
        x = torch.view_as_complex(torch.randn(half + [2], generator=g, device="cuda"))
        y = torch.empty(shape, device="cuda")
    return x, y


@pytest.mark.parametrize("desc,fuse", [("scfo8.256*4096", None), ("srfo2.60*777", None),
                                       ("srfo32x16x8*6", "off"), ("srbo32x16x8*6", "off"),
                                       ("scbo64x64*20", "on"), ("srfo48x40*9", "on")])
def test_plan_graph_replays(desc, fuse, monkeypatch):
    if fuse:
        monkeypatch.setenv("BBFFT_FUSE", fuse)
    cfg = bb.parse_fft_descriptor(desc)
    s = torch.cuda.Stream()
    plan = bb.make_plan(cfg, stream=s.cuda_stream, cache=bb.default_cache())
    eager = bb.make_plan(cfg, cache=bb.default_cache())
    x, y = _buffers(cfg, 5)
    ref = torch.empty_like(y)
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        plan.execute(x, y)  # warm-up outside the capture
    s.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        plan.execute(x, y)
    for rep in range(4):
        x.mul_(0.5 + 0.25 * rep).add_(0.125)  # eager work between replays
        y.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        eager.execute(x, ref)
        torch.cuda.synchronize()
        assert torch.equal(y, ref), (desc, rep)
