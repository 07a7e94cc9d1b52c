"""CPU: pins of oracle functions and product host code by routes independent of their code.

* slopesMaps sign, axes and normalisation: a saturated tip (tilt) puts the whole modulation ring on
  two pyramid facets, so only two quadrants are lit and the mean normalised slope is -1 up to the
  diffraction leak. Which quadrants light up follows from the mask facets (Pyramid.py:469-476), the
  tip/tilt grids (334-339) and the quadrant origins (946-961), derived in the test's comments.
* the same formula on a hand-built camera frame with constant quadrants (exact values).
* the valid-pixel threshold margin (Pyramid.py:566) of every configured sensor: the float32 frames
  of the HIP path can only reproduce the float64 selection if no pixel sits at the threshold.
* phase-screen draw order (phaseStats.py:201-271): the subharmonic draws restart RandomState(seed).
* the von Karman covariance of the extrusion (phaseStats.py:72-132) against a numerical Hankel
  transform of the phase PSD, and the extrusion operators (Atmosphere.py:335-363, 1027-1035)
  against the joint covariance they must reproduce.
"""
import math

import numpy as np
import pytest
import scipy.integrate as si
import scipy.special as sp

from oracle import pwfs as opw
from tests.helpers import config1_system


# ------------------------------------------------------------------ slopesMaps sign and normalisation
#
# Derivation (reference code only):
# * Tip = meshgrid(linspace(-pi, pi, n))[0] varies along axis 1 (columns), Tilt along axis 0 (rows)
#   (Pyramid.py:334). A phase d * Tip with d > 0 is a positive spatial frequency along columns, so
#   fft2 moves the PSF to column indices above the centre that the phasor sets (342-344): there the
#   mask coordinate x = linspace(-(1-1/N) pi, (1-1/N) pi, N) is positive (464).
# * For x > 0 the largest facet is P1 = r(x + y) when y > 0 and P4 = r(x - y) when y < 0 (469-476);
#   the mask exp(-i norma P) is a linear phase of slope -norma r 2 pi / N per pixel, which ifft2 turns
#   into a shift of +norma r pixels: P1 moves the pupil image to higher column and row indices (Q3 of
#   grabQuadrant: rows c+e.., columns c+e..), P4 to higher columns and lower rows (Q2: rows c-e-w..,
#   columns c+e..) (946-961).
# * So a large positive tip lights Q2 and Q3 only: Sx = I1 - I2 + I4 - I3 = -(I2 + I3) = -I4Q on
#   every valid pixel, and mean(Sx[valid]) / mean(I4Q[valid]) = -1 (838-853); Sy = I2 - I3 has zero
#   mean by the up-down symmetry of the pupil and of the ring. A large positive tilt lights Q3 and Q4
#   (P1 and P2 = r(-x + y)): mean Sy / norma = -1, Sx = I4 - I3 has zero mean.
# A flipped sign, swapped axes or a normalisation by the sum instead of the mean fail.

SAT = 9.0            # lambda/D: ring of radius 3 centred at 9 spans 6..12 lambda/D, all on one side


def _saturated_checks(slopes_fn):
    """Run the saturated tip and tilt through the config-1 oracle frame and ``slopes_fn``; returns
    the list of failed expectations (empty when the slopes follow the reference's convention)."""
    o = config1_system()
    g = o.geom
    w, v = g.nSubap, o.validI4Q
    bad = []
    for axis, grid, lit, dark in (('tip', o.tip, (2, 3), (1, 4)), ('tilt', o.tilt, (3, 4), (1, 2))):
        fr = o.frame(SAT * grid)
        q = {k: opw.grab_quadrant(fr, k, g)[v].sum() for k in (1, 2, 3, 4)}
        tot = sum(q.values())
        if not all(q[k] / tot > 0.49 for k in lit) or not all(q[k] / tot < 0.005 for k in dark):
            bad.append('%s lights quadrants %s' % (axis, {k: round(q[k] / tot, 4) for k in q}))
        s2d = slopes_fn(fr, v, g)
        sx, sy = s2d[:w][v].mean(), s2d[w:][v].mean()
        along, across = (sx, sy) if axis == 'tip' else (sy, sx)
        if not abs(along + 1.0) < 0.02:        # -1 up to the diffraction leak (0.94 % at 9 lambda/D)
            bad.append('%s: mean slope along the axis %.5f' % (axis, along))
        if not abs(across) < 1e-6:
            bad.append('%s: mean slope across the axis %.3e' % (axis, across))
    return bad


def _oracle_slopes(fr, v, g):
    return opw.slopes_maps(fr, v, 0, g)[0]


def test_saturated_tip_tilt_pin_slope_sign_axes_and_mean_normalisation():
    assert _saturated_checks(_oracle_slopes) == []


@pytest.mark.parametrize('mutant', ['sign', 'sum_norm', 'swap_axes'])
def test_saturated_check_rejects_planted_mistakes(mutant):
    # the check above must fail for each plausible mistake of the slopes formula, planted in a copy of
    # the oracle's output
    def mutated(fr, v, g):
        s = _oracle_slopes(fr, v, g)
        w = g.nSubap
        if mutant == 'sign':
            return -s
        if mutant == 'sum_norm':
            return s / v.sum()
        return np.concatenate((s[w:], s[:w]))
    assert _saturated_checks(mutated) != []


def hand_frame(g, values=(1.0, 2.0, 4.0, 8.0), fill=1000.0):
    """Camera frame with constant quadrants I1..I4 = values and ``fill`` elsewhere, and a valid mask
    with every third pixel invalid."""
    fr = np.full((g.cam, g.cam), fill)
    e = int(round(g.sep / 2))
    c = g.cam // 2
    w = g.nSubap
    # quadrant origins (Pyramid.py:948-961 with binning 1)
    origins = {1: (c - e - w, c - e - w), 2: (c - e - w, c + e), 3: (c + e, c + e), 4: (c + e, c - e - w)}
    for k, (r, col) in origins.items():
        fr[r:r + w, col:col + w] = values[k - 1]
    valid = (np.arange(w * w).reshape(w, w) % 3) != 0
    return fr, valid


def test_slopes_on_hand_built_frame():
    g = opw.Geometry(80, 20, 20, 10)
    fr, valid = hand_frame(g)
    s2d, s = opw.slopes_maps(fr, valid, 0, g)
    w = g.nSubap
    # I = (1, 2, 4, 8): norma = mean(I4Q[valid]) = 15, Sx = (1 - 2 + 8 - 4) / 15 = 0.2,
    # Sy = (1 - 8 + 2 - 4) / 15 = -0.6 on valid pixels, 0 elsewhere
    assert np.allclose(s2d[:w][valid], 0.2, rtol=0, atol=1e-15)
    assert np.allclose(s2d[w:][valid], -0.6, rtol=0, atol=1e-15)
    assert np.all(s2d[:w][~valid] == 0) and np.all(s2d[w:][~valid] == 0)
    assert s.size == 2 * valid.sum()


# ------------------------------------------------------------------ valid-pixel threshold margin

SENSORS = [(80, 20, 48), (160, 40, 48), (256, 64, 72), (320, 80, 72)]


@pytest.mark.parametrize('n,nSubap,T', SENSORS)
def test_valid_threshold_margin(n, nSubap, T):
    # validI4Q = I4Q >= lightRatio max(I4Q) on the float64 calibration frame (Pyramid.py:566); the HIP
    # path thresholds float32 frames (relative error ~1e-7), so every pixel must sit well away from
    # the threshold for the two selections to agree. Configured sensors: config 1, the config-3
    # sensors at 4 px per subaperture, config 2.
    pup = opw.telescope_pupil(n)
    flux = opw.flux_map(pup, opw.n_photon('I3', 8), 1e-3, 8.0)
    o = opw.PyramidOracle(n, nSubap, 3, 0.5, pup, flux, 0.8e-6, n_pix_separation=nSubap,
                          n_pix_edge=nSubap // 2, nTheta_user_defined=T, n_jobs=8)
    try:
        r = o.I4Q / o.I4Q.max()
        margin = float(np.min(np.abs(r - 0.5)))
        print('n=%d nSubap=%d T=%d: %d valid, threshold margin %.3e' % (n, nSubap, T, o.validI4Q.sum(), margin))
        assert margin > 1e-4
    finally:
        if o._pool is not None:
            o._pool.shutdown()


# ------------------------------------------------------------------ phase-screen draw order

def test_subharmonic_draws_restart_the_seeded_stream():
    # ft_sh_phase_screen makes RandomState(seed), then ft_phase_screen makes its own RandomState(seed)
    # for the FFT grid; the subharmonic levels then draw from the first instance, i.e. they reuse the
    # first numbers of the stream (phaseStats.py:176, 225, 252). Rebuild the low-order part from the
    # first 54 draws of a fresh RandomState(seed) and compare with ft_sh - ft.
    from oopao_amd.phaseStats import ft_phase_screen, ft_sh_phase_screen
    r0, L0, N, dx, seed = 0.15, 25.0, 32, 0.3, 7
    lo = ft_sh_phase_screen(r0, L0, N, dx, seed=seed) - ft_phase_screen(r0, L0, N, dx, seed=seed)
    g = np.random.RandomState(seed).normal(size=54).reshape(3, 2, 3, 3)   # per level: real, imaginary
    D = N * dx
    coords = np.arange(-N / 2, N / 2) * dx
    x, y = np.meshgrid(coords, coords)
    sh = np.zeros((N, N), complex)
    for p in range(1, 4):
        df = 1 / (3 ** p * D)
        for i in range(2):              # the first two rows and columns of the 3 x 3 grid
            for j in range(2):
                fx, fy = (j - 1) * df, (i - 1) * df
                f = math.hypot(fx, fy)
                psd = 0.0 if f == 0 else 0.023 * r0 ** (-5 / 3) / (f * f + 1 / L0 ** 2) ** (11 / 6)
                cn = (g[p - 1, 0, i, j] + 1j * g[p - 1, 1, i, j]) * math.sqrt(psd) * df
                sh += cn * np.exp(2j * np.pi * (fx * x + fy * y))
    expect = sh.real - sh.real.mean()
    assert np.max(np.abs(lo - expect)) <= 1e-12 * np.max(np.abs(expect))
    # continuing the stream after the FFT-grid draws instead gives another screen
    rs = np.random.RandomState(seed)
    rs.normal(size=2 * N * N)
    assert not np.allclose(rs.normal(size=54), g.reshape(-1))


# ------------------------------------------------------------------ von Karman covariance

def _vk_constant():
    # von Karman phase PSD K r0^(-5/3) (f^2 + 1/L0^2)^(-11/6), K = (24/5 G(6/5))^(5/6) G(11/6)^2 / (2 pi^(11/3))
    return (24 / 5 * math.gamma(6 / 5)) ** (5 / 6) * math.gamma(11 / 6) ** 2 / (2 * math.pi ** (11 / 3))


def _hankel_covariance(r, r0, L0, nz=6000):
    """B(r) = int_0^inf 2 pi f PSD(f) J0(2 pi f r) df, integrated between consecutive zeros of J0
    (Gauss-Legendre, adaptive on the first segment), the alternating tail averaged."""
    K = _vk_constant()

    def psd(f):
        return K * r0 ** (-5 / 3) * (f * f + 1 / L0 ** 2) ** (-11 / 6)
    if r == 0:
        return si.quad(lambda f: 2 * np.pi * f * psd(f), 0, np.inf, limit=500)[0]
    xg, wg = np.polynomial.legendre.leggauss(48)
    z = np.concatenate(([0.0], sp.jn_zeros(0, nz))) / (2 * np.pi * r)
    a, b = z[:-1, None], z[1:, None]
    f = 0.5 * (b - a) * xg[None, :] + 0.5 * (a + b)
    seg = 0.5 * (b - a)[:, 0] * np.sum(wg[None, :] * 2 * np.pi * f * psd(f) * sp.j0(2 * np.pi * f * r), axis=1)
    seg[0] = si.quad(lambda f: 2 * np.pi * f * psd(f) * sp.j0(2 * np.pi * f * r), 0, z[1],
                     points=[1 / L0, 10 / L0], limit=400, epsabs=0, epsrel=1e-13)[0]
    S = np.cumsum(seg)
    return 0.5 * (S[-1] + S[-2])


def test_von_karman_covariance_against_hankel_transform():
    from oopao_amd.phaseStats import makeCovarianceMatrix
    assert abs(_vk_constant() - 0.0228955) < 1e-7
    r0, L0 = 0.15, 25.0
    seps = np.array([0.0, 0.05, 0.3, 1.0, 4.0, 12.0, 40.0])
    got = makeCovarianceMatrix(np.array([0j]), seps + 0j, L0, r0)[0]
    for r, c in zip(seps, got):
        h = _hankel_covariance(r, r0, L0)
        assert abs(c / h - 1) < 1e-9, (r, c, h)
    # small separations: the structure function 2 (B(0) - B(r)) tends to Kolmogorov 6.88 (r / r0)^(5/3)
    r = 0.01
    b = makeCovarianceMatrix(np.array([0j]), np.array([0j, r + 0j]), 1e6, r0)[0]
    assert abs(2 * (b[0] - b[1]) / (6.88 * (r / r0) ** (5 / 3)) - 1) < 2e-3


# ------------------------------------------------------------------ extrusion rings and operators

def test_extrusion_ring_geometry_hand_count():
    from oopao_amd.phaseStats import extrusion_rings
    res, D = 12, 1.1
    outer, inner, innerZ, outerZ = extrusion_rings(res, 2, D)
    M = res + 2
    ii, jj = np.meshgrid(np.arange(M), np.arange(M), indexing='ij')
    dist = np.minimum(np.minimum(ii, jj), np.minimum(M - 1 - ii, M - 1 - jj))  # steps to the border
    assert np.array_equal(outer != 0, dist == 0)
    assert np.array_equal(inner != 0, (dist == 1) | (dist == 2))
    assert outer.sum() == 4 * (M - 1) == 52
    assert inner.sum() == (M - 2) ** 2 - (M - 6) ** 2 == 80
    # coordinates on a pitch D / (res - 1), row-major order (u = column, v = row)
    pitch = D / (res - 1)
    assert np.isclose(outerZ[0], 0) and np.isclose(outerZ[1], pitch)
    assert np.isclose(innerZ[0], pitch + 1j * pitch)
    assert len(innerZ) == 80 and len(outerZ) == 52


@pytest.mark.parametrize('r0', [0.15, 0.08])
def test_extrusion_operators_reproduce_the_joint_covariance(r0):
    # X = A Z + B xi with xi ~ N(0, I) independent of Z ~ N(0, c ZZt) has
    #   cov(X, Z) = A c ZZt       -> must equal c ZXt^T
    #   cov(X, X) = A c ZZt A^T + B B^T -> must equal c XXt
    # with c = (r0_def / r0)^(5/3); a transposed operand, the transient B = A of the reference's
    # buildLayer (Atmosphere.py:372) or a missing r0 scaling fails.
    from oopao_amd.phaseStats import extrusion_rings, extrusion_operators, makeCovarianceMatrix
    res, L0, r0_def = 16, 25.0, 0.15
    D = 16 * 8.0 / 12
    _, _, iz, oz = extrusion_rings(res, 2, D)
    ZZt = makeCovarianceMatrix(iz, iz, L0, r0_def)
    ZXt = makeCovarianceMatrix(iz, oz, L0, r0_def)
    XXt = makeCovarianceMatrix(oz, oz, L0, r0_def)
    A, B = extrusion_operators(ZZt, ZXt, XXt, np.linalg.pinv(ZZt), r0_def, r0)
    c = (r0_def / r0) ** (5 / 3)
    assert np.linalg.norm(A @ (c * ZZt) - c * ZXt.T) <= 1e-6 * np.linalg.norm(c * ZXt)
    joint = A @ (c * ZZt) @ A.T + B @ B.T
    assert np.linalg.norm(joint - c * XXt) <= 1e-9 * np.linalg.norm(c * XXt)
    assert np.allclose(B, np.tril(B))
    # Monte Carlo of the extrusion itself: sample covariance of the extruded ring
    rs = np.random.RandomState(3)
    w, V = np.linalg.eigh(c * ZZt)
    Z = (V * np.sqrt(np.clip(w, 0, None))) @ rs.normal(size=(len(iz), 20000))
    X = A @ Z + B @ rs.normal(size=(len(oz), 20000))
    emp = X @ X.T / 20000
    assert np.linalg.norm(emp - c * XXt) <= 0.03 * np.linalg.norm(c * XXt)
