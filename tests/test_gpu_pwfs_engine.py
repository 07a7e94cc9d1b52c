"""GPU parity of the HIP PWFS frames/slopes against the oracle (Pyramid.py:612-854).

Tolerances: the HIP path computes in float32/complex64. Frames and slopes are
compared with the fp64 oracle (the reference as shipped) at rel-L2 <= 1e-5 and
with the fp32 oracle at rel-L2 <= 1e-5.
"""
import numpy as np
import pytest

from tests.helpers import rel_l2, small_system, config1_system, config2_system, random_phase

pytestmark = pytest.mark.gpu


def _engine_for(o):
    import torch
    from oopao_amd.pwfs_engine import PWFSEngine, SlopesEngine
    from oracle.pwfs import tip_tilt
    g = o.geom
    dev = torch.device('cuda', 0)
    eng = PWFSEngine(g.n, g.N, g.cam, dev)
    tip, tilt = tip_tilt(g.n)
    eng.set_modulation(o.path, tip[0, :], tilt[:, 0])
    eng.set_mask(o.m)
    amp = torch.from_numpy(np.sqrt(o.flux / o.nTheta).astype(np.float32)).to(dev)
    return eng, amp


def _frames(eng, amp, phases):
    import torch
    ph = torch.from_numpy(np.asarray(phases, np.float32)).to(amp.device)
    out = eng.frames(ph.contiguous(), 1.0, amp)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def n160_system():
    # 40 x 40 subapertures at 4 px (N = 640: 32-lane rows pass, 16-lane columns pass)
    from oracle import pwfs as opw
    pup = opw.telescope_pupil(160)
    flux = opw.flux_map(pup, opw.n_photon('I3', 8), 1e-3, 8.0)
    return opw.PyramidOracle(160, 40, 3, 0.5, pup, flux, 0.8e-6, n_pix_separation=40, n_pix_edge=20,
                             nTheta_user_defined=16)


def n512_system():
    # 128 x 128 subapertures at 4 px (N = 2048), 4 modulation points to keep the float64 oracle short
    from oracle import pwfs as opw
    pup = opw.telescope_pupil(512)
    flux = opw.flux_map(pup, opw.n_photon('I3', 8), 1e-3, 39.0)
    return opw.PyramidOracle(512, 128, 3, 0.5, pup, flux, 0.8e-6, n_pix_separation=128, n_pix_edge=64,
                             nTheta_user_defined=4, n_jobs=4)


def _square_system(n, nsub, T):
    from oracle import pwfs as opw
    pup = opw.telescope_pupil(n)
    flux = opw.flux_map(pup, opw.n_photon('I3', 8), 1e-3, 8.0)
    return opw.PyramidOracle(n, nsub, 3, 0.5, pup, flux, 0.8e-6, n_pix_separation=nsub, n_pix_edge=nsub // 2,
                             nTheta_user_defined=T, n_jobs=4)


def n96_system():
    return _square_system(96, 24, 12)        # N = 384 (2^7 x 3), 16-lane transforms


def n128_system():
    return _square_system(128, 32, 8)        # N = 512, 32-lane transforms


def n32_system():
    return _square_system(32, 8, 8)          # N = 128, 16-lane transforms


def n64_system():
    return _square_system(64, 16, 8)         # N = 256, 32-lane transforms


def n192_system():
    return _square_system(192, 48, 8)        # N = 768 (2^8 x 3), 32-lane transforms


def n384_system():
    return _square_system(384, 96, 4)        # N = 1536 (2^9 x 3), 64-lane transforms, 4 points


def n480_system():
    # 120 x 120 subapertures at 4 px (N = 1920 = 2^7 x 15, 32-lane transforms), 4 points
    from oracle import pwfs as opw
    pup = opw.telescope_pupil(480, 0.3)
    flux = opw.flux_map(pup, opw.n_photon('I3', 8), 1e-3, 39.0)
    return opw.PyramidOracle(480, 120, 3, 0.5, pup, flux, 0.8e-6, n_pix_separation=120, n_pix_edge=60,
                             nTheta_user_defined=4, n_jobs=4)


@pytest.mark.parametrize('builder', [small_system, n32_system, n64_system, config1_system, config2_system,
                                     n96_system, n128_system, n160_system, n192_system, n384_system, n512_system,
                                     n480_system])
def test_frames_match_oracle(builder):
    o = builder()
    eng, amp = _engine_for(o)
    ph = np.stack([np.zeros((o.geom.n, o.geom.n)), random_phase(o.geom.n, 3, 1.5, o.pupil)])
    got = _frames(eng, amp, ph)
    for b in range(ph.shape[0]):
        ref = o.frame(ph[b])
        err = rel_l2(got[b], ref)
        print('frame', b, 'rel_l2 vs fp64 oracle', err)
        assert err <= 1e-5
        assert abs(got[b].sum() - ref.sum()) / ref.sum() < 1e-5


@pytest.mark.parametrize('builder', [small_system, config1_system, config2_system])
def test_slopes_match_oracle(builder):
    import torch
    from oopao_amd.pwfs_engine import SlopesEngine
    from oracle.pwfs import grab_quadrant
    o = builder()
    g = o.geom
    eng, amp = _engine_for(o)
    # quadrant origins from the reference geometry (Pyramid.py:946-961)
    e = int(np.round((g.sep / g.nSubap) * g.n / (g.n / g.nSubap) / 2 / 1))
    c = int(np.round(g.cam / 2))
    w = int(np.ceil(g.nSubap))
    quad = [c - e - w, c - e - w, c + e, c + e, c - e - w, c + e, c + e, c - e - w]
    se = SlopesEngine(amp.device, g.cam, w, quad)
    se.set_valid(o.validI4Q)
    flat = _frames(eng, amp, np.zeros((1, g.n, g.n)))
    s2d, _ = se.slopes(torch.from_numpy(flat).to(amp.device))
    se.set_reference(s2d[0])
    ph = random_phase(g.n, 5, 1.0, o.pupil)
    fr = _frames(eng, amp, ph[None])
    s2d, s = se.slopes(torch.from_numpy(fr).to(amp.device))
    torch.cuda.synchronize()
    ref2d, ref = o.measure(ph)
    err = rel_l2(s.cpu().numpy()[0], ref)
    print('slopes rel_l2 vs fp64 oracle', err, 'nsig', ref.size)
    assert s.shape[1] == o.nSignal
    assert err <= 1e-5
    o32 = builder(precision='fp32')
    err32 = rel_l2(s.cpu().numpy()[0], o32.measure(ph)[1])
    print('slopes rel_l2 vs fp32 oracle', err32)
    assert err32 <= 1e-5


def n320_t8_system():
    return small_system(n=320, nSubap=80, modulation=3, T=8)     # N = 1280, 8 points


@pytest.mark.parametrize('builder', [n96_system, n128_system, n160_system, n192_system, n320_t8_system,
                                     n384_system, n480_system])
def test_raw_and_focal_at_prime_factor_plans(builder):
    """The separable unbinned (raw_data) and focal-plane kernels against the float64 oracle, rel-L2 <=
    1e-5, at the lengths whose plans end on a prime-factor stage (N = 384 / 768 / 1536: radix 24, 1280:
    radix 20, 1920: radix 15) and at N = 512 / 640, where the focal / raw columns kernels load through
    buffer instructions and the rows pass forms the field (the kernels' compile-time variants follow the
    length)."""
    import torch
    o = builder()
    eng, amp = _engine_for(o)
    ph = random_phase(o.geom.n, 9, 1.0, o.pupil)[None]
    t = torch.from_numpy(ph.astype(np.float32)).to(amp.device).contiguous()
    raw = eng.raw(t, 1.0, amp).cpu().numpy()
    err = rel_l2(raw[0], o.raw(ph[0]))
    print('N', o.geom.N, 'raw rel_l2', err)
    assert err <= 1e-5
    crop = o.geom.N // 4
    foc = eng.focal(t, 1.0, amp, crop).cpu().numpy()
    err = rel_l2(foc[0], o.focal(ph[0], crop))
    print('N', o.geom.N, 'focal rel_l2', err)
    assert err <= 1e-5


@pytest.mark.parametrize('builder,crop', [(small_system, 20), (small_system, 64), (config1_system, 60),
                                          (config2_system, 96)])
def test_focal_camera_matches_oracle(builder, crop):
    # modulation focal-plane camera (Pyramid.py:642, 1183-1190), rel-L2 <= 1e-5 vs fp64 oracle
    import torch
    o = builder()
    eng, amp = _engine_for(o)
    ph = np.stack([random_phase(o.geom.n, 7, 1.0, o.pupil), random_phase(o.geom.n, 8, 0.5, o.pupil)])
    got = eng.focal(torch.from_numpy(ph.astype(np.float32)).to(amp.device).contiguous(), 1.0, amp, crop)
    torch.cuda.synchronize()
    got = got.cpu().numpy()
    assert got.shape == (2, crop, crop)
    for b in range(2):
        ref = o.focal(ph[b], crop)
        err = rel_l2(got[b], ref)
        print('focal', b, 'crop', crop, 'rel_l2 vs fp64 oracle', err)
        assert err <= 1e-5
