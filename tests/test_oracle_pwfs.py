"""CPU: pins of the PWFS oracle (oracle/pwfs.py) by routes independent of its code.

The reference ships no tests or golden vectors for this path, so each oracle
function is checked against: a closed form evaluated another way, a brute-force
DFT on a tiny grid, Parseval's theorem, or invariants the reference's own
construction implies (zero slopes at the calibration wavefront,
Pyramid.py:591-610).
"""
import numpy as np
import pytest

from oracle import pwfs as opw
from tests.helpers import rel_l2, small_system, random_phase


def test_photometry_and_flux():
    # Source.py:362 (I3 = 800 nm, zero point 0.58e12/368), Source.py:125
    assert opw.PHOTOMETRY['I3'][0] == 0.8e-6
    assert np.isclose(opw.n_photon('I3', 0), 0.58e12 / 368)
    assert np.isclose(opw.n_photon('I3', 5), 0.58e12 / 368 / 100)
    pup = opw.telescope_pupil(80)
    # radius (n+1)/2 on a grid of spacing n/(n-1) (Telescope.py:229-233): area pi/4 ((n+1)(n-1)/n)^2 px
    assert abs(pup.sum() - np.pi / 4 * (81 * 79 / 80) ** 2) / pup.sum() < 0.01
    assert np.array_equal(pup, pup[::-1]) and np.array_equal(pup, pup.T)


def test_geometry_config1_and_config2():
    g = opw.Geometry(80, 20, 20, 10)
    assert (g.N, g.cam, g.center) == (320, 80, 160)
    g2 = opw.Geometry(320, 80, 80, 40)
    assert (g2.N, g2.cam) == (1280, 320)
    # calibModulation clamp (Pyramid.py:292-296)
    assert g.calibModulation == 39
    with pytest.raises(ValueError):
        opw.Geometry(80, 16, 2)     # n / nSubap = 5 is odd


def test_modulation_path_counts_and_circle():
    p = opw.modulation_path(3)
    assert p.shape[0] == 4 * int(np.ceil(2 * np.pi * 3 / 4)) == 20
    p = opw.modulation_path(3, 48)
    assert p.shape == (48, 2)
    assert np.allclose(np.hypot(p[:, 0], p[:, 1]), 3.0)
    assert np.isclose(p[0, 0], 3.0) and np.isclose(p[0, 1], 0.0)


def test_mask_closed_form():
    # default pyramid (no rooftop/rotation/shift): max of the 4 facets is r(|x| + |y|)
    g = opw.Geometry(80, 20, 20, 10)
    m, mask = opw.phase_mask(g)
    N = g.N
    k = np.arange(N)
    x = -np.pi + np.pi / N + 2 * np.pi * k / N
    c = (g.n / g.nSubap) * (g.nSubap + g.sep) / 2
    closed = -c * (np.abs(x)[None, :] + np.abs(x)[:, None])
    assert np.max(np.abs(m - closed)) < 1e-9
    assert mask.dtype == np.complex64
    assert np.max(np.abs(np.abs(mask) - 1)) < 1e-6


def test_phasor_closed_form():
    N = 32
    ph = opw.phasor(N)
    w = np.exp(-1j * np.pi * (N + 1) / N)
    # phasor[y, x] = w**(x + y): ratios along both axes are constant
    assert np.allclose(ph[:, 1:] / ph[:, :-1], w)
    assert np.allclose(ph[1:, :] / ph[:-1, :], w)
    assert np.isclose(ph[0, 0], 1)


def _dft_matrix(N, sign):
    k = np.arange(N)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / N)


def test_pyramid_transform_vs_bruteforce_dft():
    o = small_system(n=8, nSubap=2, modulation=1, T=4)
    g = o.geom
    N, n, c = g.N, g.n, g.center
    amp = np.sqrt(o.flux / o.nTheta)
    ph = random_phase(n, 2, 1.0, o.pupil) + o.tt[1]
    got, _ = opw.pyramid_transform(amp, ph, o.mask, o.phasor, g)
    s = np.zeros((N, N), complex)
    s[c - n // 2:c + n // 2, c - n // 2:c + n // 2] = amp * np.exp(1j * ph)
    F = _dft_matrix(N, -1)
    Fi = _dft_matrix(N, +1) / N
    E = Fi @ ((F @ (s * o.phasor) @ F) * o.mask) @ Fi
    assert rel_l2(got, np.abs(E) ** 2) < 1e-12


def test_flux_conservation_parseval():
    # |mask| = 1 and Parseval: sum of the modulated frame = sum of the pupil intensity
    o = small_system()
    ph = random_phase(o.geom.n, 4, 2.0, o.pupil)
    raw = o.raw(ph)
    assert abs(raw.sum() / o.flux.sum() - 1) < 1e-6
    assert raw.min() >= 0


def test_zero_slopes_at_calibration_and_flux_invariance():
    o = small_system()
    s2d, s = o.measure(np.zeros((o.geom.n, o.geom.n)))
    assert np.max(np.abs(s)) == 0.0
    ph = random_phase(o.geom.n, 5, 1.0, o.pupil)
    s_a = o.measure(ph)[1]
    o2 = small_system()
    o2.flux = o2.flux * 3.7            # slopes are normalised by the mean valid flux
    assert rel_l2(o2.measure(ph)[1], s_a) < 1e-12


def test_small_tip_sign_and_symmetry():
    o = small_system()
    n = o.geom.n
    x = np.linspace(-1, 1, n)
    tip = 0.05 * np.tile(x, (n, 1)) * o.pupil           # phase ramp along x (columns)
    s2d = o.measure(tip)[0]
    w = o.geom.nSubap
    sx, sy = s2d[:w], s2d[w:]
    v = o.validI4Q
    assert abs(sx[v].mean()) > 10 * abs(sy[v].mean())
    s2d_neg = o.measure(-tip)[0]
    # small-signal linearity: opposite tips give opposite slopes
    assert rel_l2(s2d_neg[:w][v], -sx[v]) < 0.05


def test_quadrant_geometry_config1():
    g = opw.Geometry(80, 20, 20, 10)
    frame = np.arange(80 * 80, dtype=float).reshape(80, 80)
    # e = 10, c = 40, w = 20 (Pyramid.py:948-950)
    assert np.array_equal(opw.grab_quadrant(frame, 1, g), frame[10:30, 10:30])
    assert np.array_equal(opw.grab_quadrant(frame, 2, g), frame[10:30, 50:70])
    assert np.array_equal(opw.grab_quadrant(frame, 3, g), frame[50:70, 50:70])
    assert np.array_equal(opw.grab_quadrant(frame, 4, g), frame[50:70, 10:30])


def test_set_binning_sum():
    a = np.arange(64.).reshape(8, 8)
    b = opw.set_binning(a, 4)
    assert b.shape == (2, 2)
    assert b[0, 0] == a[:4, :4].sum() and b[1, 0] == a[4:, :4].sum()


def test_fp32_mode_close_to_fp64():
    o64 = small_system(precision='fp64')
    o32 = small_system(precision='fp32')
    ph = random_phase(o64.geom.n, 6, 1.0, o64.pupil)
    assert rel_l2(o32.frame(ph), o64.frame(ph)) < 1e-5


def test_focal_camera_parseval_and_centring():
    # Pyramid.py:642, 1183-1190. Parseval: the whole focal field holds the pupil flux; the
    # PSF-centring phasor puts the zero frequency at (N-1)/2, so a real field (flat phase) gives an
    # image point-symmetric about that point, and any even centred crop inherits the symmetry.
    o = small_system()
    N = o.geom.N
    ph = random_phase(o.geom.n, 6, 1.0, o.pupil)
    full = o.focal(ph, N + 1)
    assert full.shape == (N, N)
    assert abs(full.sum() / o.flux.sum() - 1) < 1e-9
    flat = o.focal(np.zeros_like(ph), N + 1)
    assert rel_l2(flat, flat[::-1, ::-1]) < 1e-12
    # the four central pixels are the maxima of the modulated ring's symmetric image
    crop = o.focal(np.zeros_like(ph), 20)
    assert crop.shape == (20, 20)
    assert rel_l2(crop, crop[::-1, ::-1]) < 1e-12
    assert rel_l2(crop, flat[N // 2 - 10:N // 2 + 10, N // 2 - 10:N // 2 + 10]) == 0
    # a single unmodulated point: the PSF peak sits on the four pixels around (N-1)/2
    o0 = small_system(modulation=0, T=None)
    f0 = o0.focal(np.zeros_like(ph), N + 1)
    peak = np.unravel_index(np.argmax(f0), f0.shape)
    assert peak[0] in (N // 2 - 1, N // 2) and peak[1] in (N // 2 - 1, N // 2)
    assert abs(f0[N // 2 - 1, N // 2 - 1] / f0[N // 2, N // 2] - 1) < 1e-9


def _frame_raw(o, ph=None):
    n = o.geom.n
    return o.raw(np.zeros((n, n)) if ph is None else ph)


def test_uniform_quadrant_shift_rolls_the_frame():
    # Pyramid.py:453-480 with sx = sy = s on all four faces: m gains -norma s x (+ norma s y), a pure
    # focal-plane ramp of 2 pi norma s / N per pixel, so the raw frame rolls by norma s pixels.
    kw = dict(n_pix_separation=4, n_pix_edge=2, nTheta_user_defined=8)
    n, nS = 16, 4
    pup = opw.telescope_pupil(n)
    flux = opw.flux_map(pup, opw.n_photon('I3', 8), 1e-3, 8.0)
    ph = random_phase(n, 3, 0.5, pup)
    base = opw.PyramidOracle(n, nS, 2, 0.1, pup, flux, 0.8e-6, **kw)
    r0 = _frame_raw(base, ph)
    ox = opw.PyramidOracle(n, nS, 2, 0.1, pup, flux, 0.8e-6, sx=(1, 1, 1, 1), validI4Q=base.validI4Q, **kw)
    oy = opw.PyramidOracle(n, nS, 2, 0.1, pup, flux, 0.8e-6, sy=(1, 1, 1, 1), validI4Q=base.validI4Q, **kw)
    norma = n // nS
    assert rel_l2(_frame_raw(ox, ph), np.roll(r0, norma, axis=1)) < 1e-6
    assert rel_l2(_frame_raw(oy, ph), np.roll(r0, -norma, axis=0)) < 1e-6


def test_spatial_filter_enters_once_or_twice():
    # Pyramid.py:633-636 and 1072: the filter is in the mask; without PSF centring it also multiplies
    # the shifted spectrum. A uniform filter 0.5 therefore scales the frame by 1/4 or by 1/16.
    n, nS = 16, 4
    pup = opw.telescope_pupil(n)
    flux = opw.flux_map(pup, opw.n_photon('I3', 8), 1e-3, 8.0)
    ph = random_phase(n, 4, 0.5, pup)
    for centring, factor in ((True, 0.25), (False, 0.0625)):
        kw = dict(n_pix_separation=4, n_pix_edge=2, nTheta_user_defined=8, psfCentering=centring)
        o = opw.PyramidOracle(n, nS, 2, 0.1, pup, flux, 0.8e-6, **kw)
        N = o.geom.N
        of = opw.PyramidOracle(n, nS, 2, 0.1, pup, flux, 0.8e-6, validI4Q=o.validI4Q,
                               spatial_filter=np.full((N, N), 0.5), **kw)
        assert rel_l2(_frame_raw(of, ph), factor * _frame_raw(o, ph)) < 1e-12


def test_weight_vector_sum():
    # Pyramid.py:753-755: raw = sum_i w_i I_i / nTheta
    o = small_system(T=8)
    ph = random_phase(o.geom.n, 5, 0.5, o.pupil)
    r0 = o.raw(ph)
    o.weight_vector = np.ones(8)
    assert rel_l2(o.raw(ph), r0 / 8) < 1e-13
    w = np.zeros(8)
    w[3] = 8.0
    o.weight_vector = w
    amp = o._amp()
    single = opw.pyramid_transform(amp, ph + o.tt[3], o.mask, o.phasor, o.geom)[0]
    assert rel_l2(o.raw(ph), single) < 1e-13


def test_old_mask_closed_form():
    # Pyramid.py:482-539: zero-mean ramps per quadrant with slopes (L +- s) norma 2 pi / h; the
    # oracle's per-pixel closed form against the drop-in's meshgrid construction.
    from types import SimpleNamespace
    from oopao_amd.Pyramid import Pyramid
    g = opw.Geometry(16, 4, 4, 2)
    N, h = g.N, g.N // 2
    L = g.nSubap + g.sep
    norma = (g.n / g.nSubap) / 4
    sx, sy = (0.5, -1.0, 0.25, 2.0), (1.5, 0.0, -0.5, 1.0)
    stub = SimpleNamespace(telescope=SimpleNamespace(resolution=16), nSubap=4)
    for centring in (True, False):
        m, mask = opw.phase_mask_old(g, centring, sx, sy)
        mp = Pyramid.get_phase_mask_old(stub, N, 4, 4, 2, centring, list(sx), list(sy))
        assert np.max(np.abs(m - mp)) < 1e-12
        assert mask.dtype == np.complex64
    m, _ = opw.phase_mask_old(g, True, sx, sy)
    q1 = -m[:h, :h]
    assert abs(q1.mean()) < 1e-12
    step = 2 * np.pi / h
    assert np.allclose(np.diff(q1, axis=1), (L + sx[0]) * norma * step)
    assert np.allclose(np.diff(q1, axis=0), (L - sy[0]) * norma * step)


def test_bin_area_pins():
    # integer factors: the reference's sum binning exactly (tools/tools.py:464-483)
    rs = np.random.RandomState(3)
    a = rs.rand(24, 24)
    assert np.allclose(opw.bin_area(a, 6), opw.set_binning(a, 4), rtol=0, atol=1e-13)
    # rational factor s = 24/5 = 4.8 (the ELT sensor's 4.8 px per subaperture): refine each raw pixel into
    # 5 x 5 equal sub-pixels, then a detector pixel is exactly 24 x 24 sub-pixels -- an independent route
    fine = np.kron(a, np.ones((5, 5))) / 25.0
    assert np.allclose(opw.bin_area(a, 5), opw.set_binning(fine, 24), rtol=0, atol=1e-13)
    # flux is conserved for any factor
    for cam in (7, 9, 11, 13):
        assert abs(opw.bin_area(a, cam).sum() - a.sum()) < 1e-12 * a.sum()


def test_fractional_geometry_config5():
    # 100 x 100 subapertures on a 480 px pupil (4.8 px each), separation 100, edge 50: N = 1920, cam = 400
    g = opw.Geometry(480, 100, 100, 50, fractional=True)
    assert (g.N, g.cam) == (1920, 400)
    assert g.N / g.cam == 4.8
    with pytest.raises(ValueError):
        opw.Geometry(480, 100, 100, 50)          # the reference's rule without the extension
