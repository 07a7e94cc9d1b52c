"""GPU: the hand-written PWFS adjoint (oopao_pwfs_adjoint) and the autograd path.

Checks, at the small system and at config 1 (80 px pupil, 20x20 subapertures, 48 points):
* the frame VJP against the float64 oracle adjoint (oracle/adjoint.py), rel-L2 <= 1e-4;
* the differentiable-slopes loss ||s(phi) - s_target||^2 over a batch of seeded screens (seeds 100..)
  against the float64 oracle gradient (slopes VJP + frame VJP), rel-L2 <= 1e-4, and against central
  finite differences of the GPU forward with eps = 1e-3 rad along random directions (rel 1e-2, the
  float32 forward's difference noise).
There is no reference gradient in OOPAO; the oracle is differentiated from its own definition and
pinned by float64 finite differences (tests/test_oracle_adjoint.py).
"""
import numpy as np
import pytest

from oracle import adjoint as oadj
from tests.helpers import rel_l2, small_system, config1_system, random_phase

pytestmark = pytest.mark.gpu


def _engine_for(o):
    import torch
    from oopao_amd.pwfs_engine import PWFSEngine
    from oracle.pwfs import tip_tilt
    g = o.geom
    dev = torch.device('cuda', 0)
    eng = PWFSEngine(g.n, g.N, g.cam, dev)
    tip, tilt = tip_tilt(g.n)
    eng.set_modulation(o.path, tip[0, :], tilt[:, 0])
    eng.set_mask(o.m)
    amp = torch.from_numpy(np.sqrt(o.flux / o.nTheta).astype(np.float32)).to(dev)
    return eng, amp


@pytest.mark.parametrize('builder', [small_system, config1_system])
def test_frame_vjp_matches_oracle(builder):
    import torch
    o = builder()
    eng, amp = _engine_for(o)
    rs = np.random.RandomState(7)
    ph = np.stack([random_phase(o.geom.n, s, 1.0, o.pupil) for s in (11, 12)])
    G = rs.randn(2, o.geom.cam, o.geom.cam)
    t = torch.from_numpy(ph.astype(np.float32)).cuda().contiguous()
    grad = eng.adjoint(t, 1.0, amp, torch.from_numpy(G.astype(np.float32)).cuda()).cpu().numpy()
    for b in range(2):
        ref = oadj.frame_vjp(o, ph[b], G[b])
        err = rel_l2(grad[b], ref)
        print('frame VJP screen', b, 'rel_l2', err)
        assert err <= 1e-4


@pytest.mark.parametrize('n', [96, 128, 192, 320, 384, 480])
def test_frame_vjp_prime_factor_plans(n):
    """The adjoint at the lengths whose plans end on a prime-factor stage: N = 384, 768, 1536 (radix 24,
    8 x 3), N = 1280 (radix 20, 4 x 5) and N = 1920 (radix 15, 3 x 5), and at N = 512, where the adjoint's
    columns kernel loads and stores through buffer instructions (as at 768, 1280 and 1920); few points
    (the plan depends on the length only)."""
    import torch
    o = small_system(n=n, nSubap=n // 4, modulation=3, T=8)
    assert o.geom.N == 4 * n
    eng, amp = _engine_for(o)
    ph = random_phase(o.geom.n, 21, 1.0, o.pupil)[None]
    G = np.random.RandomState(8).randn(1, o.geom.cam, o.geom.cam)
    t = torch.from_numpy(ph.astype(np.float32)).cuda().contiguous()
    grad = eng.adjoint(t, 1.0, amp, torch.from_numpy(G.astype(np.float32)).cuda()).cpu().numpy()
    err = rel_l2(grad[0], oadj.frame_vjp(o, ph[0], G[0]))
    print('N', o.geom.N, 'frame VJP rel_l2', err)
    assert err <= 1e-4


def test_slopes_loss_gradient_config1():
    import torch
    from oopao_amd.differentiable import pwfs_frames, slopes_maps
    o = config1_system()
    eng, amp = _engine_for(o)
    g = o.geom
    e = int(np.round((g.sep / g.nSubap) * g.n / (g.n / g.nSubap) / 2))
    c = int(np.round(g.cam / 2))
    w = int(np.ceil(g.nSubap))
    quad = [c - e - w, c - e - w, c + e, c + e, c - e - w, c + e, c + e, c - e - w]
    valid = torch.from_numpy(o.validI4Q).cuda()
    ref2d = torch.from_numpy(o.ref2d.astype(np.float32)).cuda()
    B = 64
    ph = np.stack([random_phase(g.n, 100 + k, 0.5, o.pupil) for k in range(B)])
    target = o.measure(random_phase(g.n, 999, 0.5, o.pupil))[1]
    tgt = torch.from_numpy(target.astype(np.float32)).cuda()

    def loss_fn(phase):
        # slopes in float32 from the GPU path; the loss itself is reduced in float64 so that a
        # central difference at eps = 1e-3 is not swamped by the rounding of a float32 sum
        frames = pwfs_frames(eng, phase, amp, 1.0)
        s = slopes_maps(frames, valid, quad, w, ref2d, 1.0)[1]
        return ((s.double() - tgt.double()[None]) ** 2).sum()

    phase = torch.from_numpy(ph.astype(np.float32)).cuda().requires_grad_(True)
    L = loss_fn(phase)
    L.backward()
    grad = phase.grad.detach().cpu().numpy()
    # float64 oracle gradient on a subset of the screens
    for b in (0, 17, 63):
        s = o.measure(ph[b])[1]
        dsig = 2 * (s - target)
        gF = oadj.slopes_vjp(o, o.frame(ph[b]), dsig)
        ref = oadj.frame_vjp(o, ph[b], gF)
        err = rel_l2(grad[b], ref)
        print('loss gradient screen', b, 'rel_l2 vs float64 oracle', err)
        assert err <= 1e-4
    # central finite differences of the GPU forward, eps = 1e-3 rad, random directions
    rs = np.random.RandomState(1)
    with torch.no_grad():
        for k in range(3):
            d = torch.from_numpy((rs.randn(*ph.shape) * o.pupil).astype(np.float32)).cuda()
            eps = 1e-3
            fd = (loss_fn(phase + eps * d) - loss_fn(phase - eps * d)).item() / (2 * eps)
            an = float((phase.grad * d).sum().item())
            print('FD', fd, 'adjoint', an)
            assert abs(fd - an) <= 1e-2 * abs(an)


@pytest.mark.parametrize('builder', [small_system, config1_system])
def test_banded_adjoint_equals_full(builder):
    """A cotangent that is zero outside two detector-column bands (as for any loss of the slopesMaps
    signal): the banded adjoint (only those columns transformed) gives the full adjoint's gradient bit
    for bit, and the banded forward gives the full frame on those columns bit for bit."""
    import torch
    o = builder()
    eng, amp = _engine_for(o)
    cam = o.geom.cam
    bands = [(cam // 8, 3 * cam // 8), (5 * cam // 8, 7 * cam // 8)]
    rs = np.random.RandomState(8)
    ph = np.stack([random_phase(o.geom.n, s, 1.0, o.pupil) for s in (21, 22, 23)])
    G = rs.randn(3, cam, cam)
    cols = np.zeros(cam, bool)
    for a, b in bands:
        cols[a:b] = True
    G[:, :, ~cols] = 0.0
    t = torch.from_numpy(ph.astype(np.float32)).cuda().contiguous()
    g = torch.from_numpy(G.astype(np.float32)).cuda()
    full = eng.adjoint(t, 1.0, amp, g)
    banded = eng.adjoint(t, 1.0, amp, g, bands=bands)
    print('banded vs full adjoint: max |diff|', float((banded - full).abs().max()))
    assert torch.equal(banded, full)
    f_full = eng.frames(t, 1.0, amp)
    f_band = eng.frames(t, 1.0, amp, bands=bands)
    assert torch.equal(f_band[:, :, torch.from_numpy(cols).cuda()], f_full[:, :, torch.from_numpy(cols).cuda()])
    assert float(f_band[:, :, torch.from_numpy(~cols).cuda()].abs().max()) == 0.0
