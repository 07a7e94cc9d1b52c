"""GPU: the general-path PWFS adjoint (oopao_pwfs_generic_adjoint) and the autograd path through it.

OOPAO has no gradients; the forward differentiated is Pyramid.pyramid_transform / wfs_measure /
signalProcessing (Pyramid.py:612-757, 835-854) and the Bi-O Edge transform (BioEdge.py:510-544).
* frame VJP against the float64 oracle adjoint (oracle/adjoint.py frame_vjp, pinned by float64 finite
  differences in tests/test_oracle_adjoint.py), rel-L2 <= 1e-4, for sensors outside the separable path:
  OOPAO's default geometry (n_pix_separation = 2: N = 264 = 8 x 3 x 11 with 6 x 6 binning at n = 120),
  N = 56 (radix 7), a rooftop and a rotated pyramid; the opd scale is checked with scale = 2.5;
* the drop-in Pyramid at its default geometry (general path): gradient of <c, s(phi)> through
  pwfs_signal against the oracle slopes VJP + frame VJP, rel-L2 <= 1e-4;
* the drop-in Pyramid with the detector's extra binning (2 x 2, a torch op in the autograd graph):
  central differences of the GPU forward at eps = 1e-3 rad along smooth directions (white-noise ones
  barely move the slopes, and the float32 forward's rounding then dominates the difference), relative
  1e-2;
* Bi-O Edge (four knife-edge masks on the general path): directional derivatives of <c, s(phi)>
  against central differences of the float64 Bi-O Edge oracle (oracle/bioedge.py) at eps = 1e-4 rad,
  relative 1e-3.
"""
import numpy as np
import pytest

from oracle import adjoint as oadj
from oracle import pwfs as opw
from tests.helpers import rel_l2, random_phase

pytestmark = pytest.mark.gpu


def _oracle(n, nSubap, sep, edge, modulation=3, T=None, **kw):
    pup = opw.telescope_pupil(n)
    flux = opw.flux_map(pup, opw.n_photon('I3', 8), 1e-3, 8.0)
    return opw.PyramidOracle(n, nSubap, modulation, 0.5, pup, flux, 0.8e-6, n_pix_separation=sep,
                             n_pix_edge=edge, nTheta_user_defined=T, **kw)


def _engine(o):
    import torch
    from oopao_amd.pwfs_engine import PWFSEngine
    g = o.geom
    dev = torch.device('cuda', 0)
    eng = PWFSEngine(g.n, g.N, g.cam, dev)
    tip, tilt = opw.tip_tilt(g.n)
    eng.set_modulation(o.path, tip[0, :], tilt[:, 0], o.psfCentering)
    eng.set_mask(o.m, o.mask)
    amp = torch.from_numpy(np.sqrt(o.flux / o.nTheta).astype(np.float32)).to(dev)
    return eng, amp


CASES = {'tutorial_N264': ((120, 20, 2, 1), {}), 'N56_radix7': ((24, 12, 2, 1), {}),
         'rooftop': ((80, 20, 20, 10), dict(T=16, rooftop=0.5)),
         'rotated': ((80, 20, 20, 10), dict(T=16, theta_rotation=0.1))}


@pytest.mark.parametrize('case', sorted(CASES))
def test_generic_frame_vjp_matches_oracle(case):
    import torch
    args, kw = CASES[case]
    o = _oracle(*args, **kw)
    eng, amp = _engine(o)
    g = o.geom
    print('generic adjoint', case, 'N', g.N, 'cam', g.cam, 'T', o.nTheta, flush=True)
    assert eng.mode == 'generic'
    rs = np.random.RandomState(7)
    ph = np.stack([random_phase(g.n, s, 1.0, o.pupil) for s in (11, 12)])
    G = rs.randn(2, g.cam, g.cam)
    scale = 2.5
    opd = torch.from_numpy((ph / scale).astype(np.float32)).cuda().contiguous()
    grad = eng.adjoint(opd, scale, amp, torch.from_numpy(G.astype(np.float32)).cuda()).cpu().numpy()
    for b in range(2):
        ref = scale * oadj.frame_vjp(o, ph[b], G[b])
        err = rel_l2(grad[b], ref)
        print('  screen', b, 'rel_l2 vs float64 oracle %.2e' % err, flush=True)
        assert err <= 1e-4


def _dropin(n, nSubap, **kw):
    from oopao_amd import Telescope, Source, Pyramid
    tel = Telescope(n, 8.0, 1e-3)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    wfs = Pyramid(nSubap=nSubap, telescope=tel, modulation=3, lightRatio=0.5, **kw)
    return tel, ngs, wfs


def test_dropin_default_geometry_slopes_gradient():
    import torch
    from oopao_amd.differentiable import pwfs_signal
    tel, ngs, wfs = _dropin(80, 20, nTheta_user_defined=16)
    o = _oracle(80, 20, 2, 1, T=16)
    print('drop-in default geometry: N', wfs.engine.N, 'mode', wfs.engine.mode, flush=True)
    assert wfs.engine.mode == 'generic' and wfs.engine.N == o.geom.N and wfs.cam.resolution == o.geom.cam
    o.validI4Q = np.asarray(wfs.validI4Q, bool)
    rs = np.random.RandomState(3)
    B = 3
    ph = np.stack([random_phase(80, 200 + k, 0.5, o.pupil) for k in range(B)])
    c = rs.randn(B, int(wfs.nSignal))
    phase = torch.from_numpy(ph.astype(np.float32)).cuda().requires_grad_(True)
    s = pwfs_signal(wfs, phase)
    (s.double() * torch.from_numpy(c).cuda()).sum().backward()
    grad = phase.grad.cpu().numpy()
    for b in range(B):
        gF = oadj.slopes_vjp(o, o.frame(ph[b]), c[b])
        ref = oadj.frame_vjp(o, ph[b], gF)
        err = rel_l2(grad[b], ref)
        print('  screen', b, 'rel_l2 vs float64 oracle %.2e' % err, flush=True)
        assert err <= 1e-4


def test_dropin_binning_gradient_finite_differences():
    import torch
    from oopao_amd.differentiable import pwfs_signal
    tel, ngs, wfs = _dropin(80, 20, n_pix_separation=20, n_pix_edge=10, nTheta_user_defined=16, binning=2)
    print('drop-in binning 2: mode', wfs.engine.mode, 'nSignal', wfs.nSignal, flush=True)
    rs = np.random.RandomState(4)
    ph = random_phase(80, 300, 0.5, tel.pupil)[None]
    c = torch.from_numpy(rs.randn(int(wfs.nSignal))).cuda()

    def loss(p):
        return (pwfs_signal(wfs, p).double()[0] * c).sum()

    phase = torch.from_numpy(ph.astype(np.float32)).cuda().requires_grad_(True)
    loss(phase).backward()
    with torch.no_grad():
        for k in range(3):
            d = torch.from_numpy(random_phase(80, 500 + k, 1.0, tel.pupil)[None].astype(np.float32)).cuda()
            eps = 1e-3
            fd = (loss(phase + eps * d) - loss(phase - eps * d)).item() / (2 * eps)
            an = float((phase.grad * d).sum().item())
            print('  FD %.6e adjoint %.6e' % (fd, an), flush=True)
            assert abs(fd - an) <= 1e-2 * abs(an)


def test_bioedge_slopes_gradient_matches_oracle_differences():
    import torch
    from oopao_amd import Telescope, Source
    from oopao_amd.BioEdge import BioEdge
    from oopao_amd.differentiable import pwfs_signal
    from oracle.bioedge import BioEdgeOracle
    tel = Telescope(32, 8, 1e-3)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    kw = dict(modulation=2, grey_width=1, lightRatio=0.1, n_pix_separation=4, nTheta_user_defined=16,
              postProcessing='slopesMaps')
    wfs = BioEdge(nSubap=8, telescope=tel, **kw)
    flux = opw.flux_map(tel.pupil, ngs.nPhoton, tel.samplingTime, tel.D)
    o = BioEdgeOracle(tel.resolution, wfs.nSubap, wfs.modulation, wfs.lightRatio, tel.pupil, flux,
                      incidence_norma=ngs.nPhoton * tel.samplingTime * (tel.D / wfs.nSubap) ** 2,
                      postProcessing='slopesMaps', n_pix_separation=4, grey_width=1, nTheta_user_defined=16)
    assert np.array_equal(np.asarray(wfs.validSignal), o.validSignal)
    print('Bi-O Edge gradient: N', wfs.engine.N, 'nSignal', wfs.nSignal, flush=True)
    rs = np.random.RandomState(5)
    ph0 = random_phase(32, 400, 0.8, tel.pupil)
    c = rs.randn(int(wfs.nSignal))
    phase = torch.from_numpy(ph0[None].astype(np.float32)).cuda().requires_grad_(True)
    s = pwfs_signal(wfs, phase)
    (s.double()[0] * torch.from_numpy(c).cuda()).sum().backward()
    grad = phase.grad.cpu().numpy()[0]
    eps = 1e-4
    for k in range(3):
        d = rs.randn(32, 32) * tel.pupil
        fd = (c @ o.measure(ph0 + eps * d)[1] - c @ o.measure(ph0 - eps * d)[1]) / (2 * eps)
        an = float(np.sum(grad * d))
        print('  oracle FD %.6e adjoint %.6e' % (fd, an), flush=True)
        assert abs(fd - an) <= 1e-3 * abs(fd)
