"""GPU: the first PWFS call of a new engine made under graph capture.

An engine's construction prepares everything its calls would otherwise set up on first use (the
twiddle table of its length and the occupancy of its rows-pass kernels), so its very first frame call
can be recorded into a HIP graph: the replays, with eager work and synchronisations between them, give
the eager frames bit for bit, slopes-only and whole.
"""
import numpy as np
import pytest

from tests.helpers import small_system, random_phase

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('whole', [False, True])
def test_first_frame_call_under_capture(whole):
    import torch
    from oopao_amd.pwfs_engine import PWFSEngine
    from oracle.pwfs import tip_tilt
    # N = 384 (96 px pupil, 24 subapertures): a fast-path length the other graph tests do not use
    o = small_system(n=96, nSubap=24, modulation=3, T=36)
    g = o.geom
    assert g.N == 384
    dev = torch.device('cuda', 0)
    eng = PWFSEngine(g.n, g.N, g.cam, dev)          # construction prepares the length's tables
    tip, tilt = tip_tilt(g.n)
    eng.set_modulation(o.path, tip[0, :], tilt[:, 0])
    eng.set_mask(o.m)
    amp = torch.from_numpy(np.sqrt(o.flux / o.nTheta).astype(np.float32)).to(dev)
    ph = np.stack([random_phase(g.n, 70 + k, 1.0, o.pupil) for k in range(2)])
    opd = torch.from_numpy(ph.astype(np.float32)).to(dev).contiguous()
    cam = g.cam
    bands = None if whole else [(cam // 8, 3 * cam // 8), (5 * cam // 8, 7 * cam // 8)]
    out = torch.zeros((2, cam, cam), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    stream = torch.cuda.Stream(dev)
    stream.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.graph(graph, stream=stream):
        eng.frames(opd, 1.0, amp, out=out, bands=bands)   # the engine's first frame call
    eager = eng.frames(opd, 1.0, amp, bands=bands)
    torch.cuda.synchronize()
    for rep in range(3):
        out.zero_()
        graph.replay()
        torch.cuda.synchronize()
        print('replay', rep, 'max |diff| vs eager', float((out - eager).abs().max()), flush=True)
        assert torch.equal(out, eager)
        _ = eng.frames(opd, 1.0, amp, bands=bands)         # eager work between the replays
        torch.cuda.synchronize()
    assert float(eager.abs().max()) > 0
