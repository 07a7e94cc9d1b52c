"""GPU parity of the general PWFS path (any N, any mask, any binning) against the oracle.

Reference: Pyramid.pyramid_transform / wfs_measure (Pyramid.py:612-757), the
mask options (rooftop, rotation, per-quadrant shifts, old mask: 453-539), the
spatial filter (633-636, 1055-1085), weight_vector (750-755), raw_data and the
focal-plane camera (642, 1183-1190). Geometries cover the radix paths of the
runtime-length FFT: N = 264 (8*3*11, OOPAO's tutorial sensor with
n_pix_separation = 2), 56 (7), 104 (13), 92 (4*23, direct odd radix), 496 (16*31).
Tolerance: rel-L2 <= 1e-5 against the float64 oracle (float32 GPU arithmetic).
"""
import numpy as np
import pytest

from oracle import pwfs as opw
from tests.helpers import rel_l2, random_phase

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _oracle(n, nSubap, sep, edge, modulation=3, T=None, **kw):
    pup = opw.telescope_pupil(n)
    flux = opw.flux_map(pup, opw.n_photon('I3', 8), 1e-3, 8.0)
    return opw.PyramidOracle(n, nSubap, modulation, 0.5, pup, flux, 0.8e-6, n_pix_separation=sep,
                             n_pix_edge=edge, nTheta_user_defined=T, **kw)


def _engine(o, factor=None, general=False):
    import torch
    from oopao_amd.pwfs_engine import PWFSEngine
    g = o.geom
    dev = torch.device('cuda', 0)
    eng = PWFSEngine(g.n, g.N, g.cam, dev)
    tip, tilt = opw.tip_tilt(g.n)
    eng.set_modulation(o.path, tip[0, :], tilt[:, 0], o.psfCentering)
    eng.set_mask(o.m, o.mask, factor=factor, general=general)
    amp = torch.from_numpy(np.sqrt(o.flux / o.nTheta).astype(np.float32)).to(dev)
    return eng, amp


def _phases(o, seeds=(11, 12)):
    n = o.geom.n
    return np.stack([random_phase(n, s, 1.0, o.pupil) for s in seeds])


def _frames(eng, amp, ph, **kw):
    import torch
    t = torch.from_numpy(ph.astype(np.float32)).to(amp.device).contiguous()
    out = eng.frames(t, 1.0, amp, **kw)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _check_frames(o, eng, amp, ph, label):
    got = _frames(eng, amp, ph)
    for b in range(ph.shape[0]):
        err = rel_l2(got[b], o.frame(ph[b]))
        print(label, 'N', o.geom.N, 'screen', b, 'rel_l2', err)
        assert err <= TOL


@pytest.mark.parametrize('geom', [(120, 20, 2, 1), (24, 12, 2, 1), (48, 24, 2, 1), (40, 20, 3, 1.5),
                                  (240, 60, 2, 1), (160, 80, 2, 1), (208, 104, 2, 1)])
def test_generic_geometries_frames(geom):
    # N = 264, 56, 104, 92, 496, 328 (= 8 x 41) and 424 (= 8 x 53, the LDS direct-DFT stage)
    o = _oracle(*geom)
    eng, amp = _engine(o)
    assert eng.mode == 'generic'
    _check_frames(o, eng, amp, _phases(o), 'geometry %s' % (geom,))


@pytest.mark.parametrize('rooftop,theta', [(0.5, 0.0), (0.0, 0.1), (0.3, -0.05)])
def test_non_separable_masks(rooftop, theta):
    o = _oracle(80, 20, 20, 10, T=48, rooftop=rooftop, theta_rotation=theta)
    eng, amp = _engine(o)
    assert eng.mode == 'generic'
    _check_frames(o, eng, amp, _phases(o), 'rooftop %g theta %g' % (rooftop, theta))


def test_per_quadrant_shifts_and_old_mask():
    sx, sy = (0.5, -1.0, 0.0, 1.0), (0.0, 0.5, -0.5, 0.0)
    o = _oracle(80, 20, 20, 10, T=48, sx=sx, sy=sy)
    eng, amp = _engine(o)
    assert eng.mode == 'generic'
    _check_frames(o, eng, amp, _phases(o), 'shifts')
    for centring in (True, False):
        o = _oracle(80, 20, 20, 10, T=48, old_mask=True, sx=sx, sy=sy, psfCentering=centring)
        eng, amp = _engine(o)
        _check_frames(o, eng, amp, _phases(o), 'old mask centring %s' % centring)


@pytest.mark.parametrize('geom', [(80, 20, 20, 10), (120, 20, 2, 1)])
def test_without_psf_centring(geom):
    o = _oracle(*geom, psfCentering=False)
    eng, amp = _engine(o)
    _check_frames(o, eng, amp, _phases(o), 'no centring, mode %s' % eng.mode)


@pytest.mark.parametrize('centring', [True, False])
def test_spatial_filter(centring):
    # circular focal-plane stop of radius N/6 (SpatialFilter-like), applied as in Pyramid.py:633-636
    n, nS = 80, 20
    g = opw.Geometry(n, nS, 20, 10)
    N = g.N
    y, x = np.mgrid[0:N, 0:N] - (N - 1) / 2
    sf = (x ** 2 + y ** 2 <= (N / 6) ** 2).astype(np.float64) * (1.0 + 0.2 * np.cos(x / 7.0))
    o = _oracle(n, nS, 20, 10, T=48, psfCentering=centring, spatial_filter=sf)
    eng, amp = _engine(o, factor=None if centring else sf, general=True)
    _check_frames(o, eng, amp, _phases(o), 'spatial filter centring %s' % centring)


def test_weight_vector():
    o = _oracle(80, 20, 20, 10, T=48)
    w = np.random.RandomState(3).uniform(0.2, 2.0, 48)
    eng, amp = _engine(o)
    eng.set_weights(w)
    o.weight_vector = w
    ph = _phases(o)
    got = _frames(eng, amp, ph, weighted=True)
    for b in range(2):
        err = rel_l2(got[b], o.frame(ph[b]))
        print('weighted frame', b, err)
        assert err <= TOL


@pytest.mark.parametrize('geom', [(80, 20, 20, 10), (120, 20, 2, 1)])
def test_raw_frame(geom):
    import torch
    o = _oracle(*geom)
    eng, amp = _engine(o)
    ph = _phases(o)
    raw = eng.raw(torch.from_numpy(ph.astype(np.float32)).to(amp.device).contiguous(), 1.0, amp)
    torch.cuda.synchronize()
    raw = raw.cpu().numpy()
    for b in range(2):
        err = rel_l2(raw[b], o.raw(ph[b]))
        print('raw', geom, b, err)
        assert err <= TOL


@pytest.mark.parametrize('crop', [40, 120])
def test_generic_focal_camera(crop):
    import torch
    o = _oracle(120, 20, 2, 1)
    eng, amp = _engine(o)
    ph = _phases(o)
    got = eng.focal(torch.from_numpy(ph.astype(np.float32)).to(amp.device).contiguous(), 1.0, amp, crop)
    torch.cuda.synchronize()
    got = got.cpu().numpy()
    for b in range(2):
        err = rel_l2(got[b], o.focal(ph[b], crop))
        print('generic focal', crop, b, err)
        assert err <= TOL


def test_dropin_tutorial_sensor_slopes():
    # OOPAO's tutorial sensor: 6 px per subaperture, n_pix_separation 2, edge 1 (N = 264, 6 x 6 binning)
    from oopao_amd import Telescope, Source, Pyramid
    tel = Telescope(120, 8, 1e-3, centralObstruction=0)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    wfs = Pyramid(20, tel, 3, 0.5, n_pix_separation=2, n_pix_edge=1)
    o = _oracle(120, 20, 2, 1)
    assert wfs.resolution == o.geom.N == 264
    assert wfs.engine.mode == 'generic'
    assert np.array_equal(wfs.validI4Q, o.validI4Q)
    ph = random_phase(120, 21, 0.8, o.pupil)
    ngs ** tel
    tel.OPD = ph * o.wavelength / (2 * np.pi)
    wfs.wfs_measure()
    err = rel_l2(wfs.signal, o.measure(ph)[1])
    print('tutorial sensor slopes rel_l2', err)
    assert err <= TOL
    err = rel_l2(wfs.raw_data, o.raw(ph))
    print('tutorial sensor raw_data rel_l2', err)
    assert err <= TOL
