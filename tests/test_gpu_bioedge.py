"""GPU: the Bi-O Edge drop-in (oopao_amd/BioEdge.py) against the float64 oracle (oracle/bioedge.py).

Reference: BioEdge.py (masks 331-390, transform 510-544, measurement 559-727, processing 729-820,
quadrants 840-853, detector 1033-1055); the tutorial's sensor (tutorials/AO_closed_loop_BioEdge_WFS.py:
144-151: modulation 0, grey width 3, separation 0, fullFrame_incidence_flux) at a smaller pupil.
Tolerance: float32 GPU vs float64 oracle, rel-L2 <= 1e-5 on frames and signals.
"""
import numpy as np
import pytest

from oracle import pwfs as opw
from oracle.bioedge import BioEdgeOracle
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def _setup(n=32, nSubap=8):
    from oopao_amd import Telescope, Source
    tel = Telescope(n, 8, 1e-3)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    return tel, ngs


def _opd(tel, seed, rms=60e-9):
    rs = np.random.RandomState(seed)
    o = rs.randn(tel.resolution, tel.resolution)
    from scipy.ndimage import uniform_filter
    o = uniform_filter(o, 5)
    return o / o.std() * rms


def _oracle(tel, ngs, wfs, **kw):
    flux = opw.flux_map(tel.pupil, ngs.nPhoton, tel.samplingTime, tel.D)
    return BioEdgeOracle(tel.resolution, wfs.nSubap, wfs.modulation, wfs.lightRatio, tel.pupil, flux,
                         incidence_norma=ngs.nPhoton * tel.samplingTime * (tel.D / wfs.nSubap) ** 2, **kw)


@pytest.mark.parametrize('case', ['tutorial', 'modulated_slopes'])
def test_bioedge_signals(case):
    from oopao_amd.BioEdge import BioEdge
    print('bioedge', case, flush=True)
    tel, ngs = _setup()
    if case == 'tutorial':
        kw = dict(modulation=0, grey_width=3, grey_length=False, lightRatio=0.01, n_pix_separation=0,
                  postProcessing='fullFrame_incidence_flux')
        okw = dict(postProcessing='fullFrame_incidence_flux', n_pix_separation=0, grey_width=3)
    else:
        kw = dict(modulation=2, grey_width=1, lightRatio=0.1, n_pix_separation=4, nTheta_user_defined=16,
                  postProcessing='slopesMaps')
        okw = dict(postProcessing='slopesMaps', n_pix_separation=4, grey_width=1, nTheta_user_defined=16)
    wfs = BioEdge(nSubap=8, telescope=tel, **kw)
    o = _oracle(tel, ngs, wfs, **okw)
    assert wfs.cam.resolution == o.g.cam and wfs.resolution == o.g.N
    assert np.array_equal(np.asarray(wfs.validSignal), o.validSignal)
    opd = _opd(tel, 3)
    tel.OPD_no_pupil = opd
    ngs * tel * wfs
    phase = opd * tel.pupil * 2 * np.pi / ngs.wavelength
    f_ref = o.frame(phase)
    s2d_ref, s_ref = o.measure(phase)
    ef, es = rel_l2(wfs.cam.frame, f_ref), rel_l2(wfs.signal, s_ref)
    e2 = rel_l2(wfs.signal_2D, s2d_ref)
    print('frame %.2e signal %.2e signal_2D %.2e' % (ef, es, e2))
    assert ef < 1e-5 and es < 1e-5 and e2 < 1e-5
    # raw 2N x 2N frame and the focal-plane camera
    assert rel_l2(wfs.raw_data, o.raw(phase)) < 1e-5
    wfs.focal_plane_camera.resolution = wfs.resolution
    wfs * wfs.focal_plane_camera
    assert rel_l2(wfs.focal_plane_camera.frame, o.focal(phase, wfs.resolution)) < 1e-5


def test_bioedge_grey_length_and_shifts():
    from oopao_amd.BioEdge import BioEdge
    print('bioedge grey_length + shifts', flush=True)
    tel, ngs = _setup()
    wfs = BioEdge(nSubap=8, telescope=tel, modulation=1, lightRatio=0.1, n_pix_separation=4, grey_width=1,
                  grey_length=2, nTheta_user_defined=8)
    sx, sy = [1, 0, 0, 0], [0, 0, 1, 0]
    wfs.apply_shift_wfs(sx=sx, sy=sy)
    o = _oracle(tel, ngs, wfs, postProcessing='slopesMaps', n_pix_separation=4, grey_width=1, grey_length=2,
                nTheta_user_defined=8, sx=sx, sy=sy)
    opd = _opd(tel, 4)
    tel.OPD_no_pupil = opd
    ngs * tel * wfs
    phase = opd * tel.pupil * 2 * np.pi / ngs.wavelength
    err = rel_l2(wfs.cam.frame, o.frame(phase))
    print('frame rel-L2 %.2e' % err)
    assert err < 1e-5


def test_bioedge_workspace_failure_halves_the_batch(monkeypatch):
    # a workspace above the host-side cap makes the Bi-O Edge measure fewer screens per call
    import torch
    from oopao_amd.BioEdge import BioEdge
    tel, ngs = _setup()
    kw = dict(modulation=2, grey_width=1, lightRatio=0.1, n_pix_separation=4, nTheta_user_defined=16,
              postProcessing='slopesMaps')
    maps = torch.from_numpy(np.stack([_opd(tel, 10 + k) for k in range(6)]).astype(np.float32)).cuda()
    wfs = BioEdge(nSubap=8, telescope=tel, **kw)
    ngs * tel * wfs
    ref = wfs.frames_dev(maps * tel._pupil_dev)
    wfs2 = BioEdge(nSubap=8, telescope=tel, **kw)
    ngs * tel * wfs2
    per = wfs2.engine.bytes_per_screen()
    monkeypatch.setenv('OOPAO_AMD_MAX_WORKSPACE_BYTES', str(int(2.5 * per)))
    got = wfs2.frames_dev(maps * tel._pupil_dev)
    monkeypatch.delenv('OOPAO_AMD_MAX_WORKSPACE_BYTES')
    assert wfs2.max_batch_bytes <= 2 * per
    assert rel_l2(got.cpu().numpy(), ref.cpu().numpy()) <= 1e-6


def test_bioedge_rank_one_masks_take_the_separable_kernels():
    # N = 4 n with 4 x 4 binning (n = 80, 20 subapertures, separation 20): each knife-edge mask (an
    # amplitude along one axis times a tip / tilt phase) factorises as p[k] q[l] and runs on the two-pass
    # kernels; frames, signals, the raw frame and the adjoint against the float64 oracle. The grey-length
    # masks (quadrant blocks) are not of rank one and stay on the general path.
    import torch
    from oopao_amd.BioEdge import BioEdge
    from oopao_amd.differentiable import pwfs_signal
    print('bioedge rank-one masks', flush=True)
    tel, ngs = _setup(80, 20)
    kw = dict(modulation=3, grey_width=1, lightRatio=0.1, n_pix_separation=20, nTheta_user_defined=16,
              postProcessing='slopesMaps')
    wfs = BioEdge(nSubap=20, telescope=tel, **kw)
    assert [e.mode for e in wfs.engines] == ['fast'] * 4
    o = _oracle(tel, ngs, wfs, postProcessing='slopesMaps', n_pix_separation=20, grey_width=1,
                nTheta_user_defined=16)
    assert np.array_equal(np.asarray(wfs.validSignal), o.validSignal)
    opd = _opd(tel, 5)
    tel.OPD_no_pupil = opd
    ngs * tel * wfs
    phase = opd * tel.pupil * 2 * np.pi / ngs.wavelength
    ef, es = rel_l2(wfs.cam.frame, o.frame(phase)), rel_l2(wfs.signal, o.measure(phase)[1])
    er = rel_l2(wfs.raw_data, o.raw(phase))
    print('fast-path Bi-O Edge: frame %.2e signal %.2e raw %.2e' % (ef, es, er), flush=True)
    assert ef < 1e-5 and es < 1e-5 and er < 1e-5
    # the separable adjoint: directional derivatives of <c, s(phi)> against float64 oracle differences
    rs = np.random.RandomState(6)
    c = rs.randn(int(wfs.nSignal))
    ph = torch.from_numpy(phase[None].astype(np.float32)).cuda().requires_grad_(True)
    (pwfs_signal(wfs, ph).double()[0] * torch.from_numpy(c).cuda()).sum().backward()
    grad = ph.grad.cpu().numpy()[0]
    for k in range(2):
        d = rs.randn(80, 80) * tel.pupil
        fd = (c @ o.measure(phase + 1e-4 * d)[1] - c @ o.measure(phase - 1e-4 * d)[1]) / 2e-4
        an = float(np.sum(grad * d))
        print('  oracle FD %.6e adjoint %.6e' % (fd, an), flush=True)
        assert abs(fd - an) <= 1e-3 * abs(fd)
    wfs_gl = BioEdge(nSubap=20, telescope=tel, grey_length=2, **kw)
    assert [e.mode for e in wfs_gl.engines] == ['generic'] * 4
