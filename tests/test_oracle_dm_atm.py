"""CPU: pins of the DM and atmosphere oracles (oracle/dm.py, oracle/atmosphere.py) and of the
host-side phase-screen code (oopao_amd/phaseStats.py)."""
import math

import numpy as np
import pytest
from numpy.random import RandomState

from oracle.dm import DMOracle
from oracle import atmosphere as oat
from tests.helpers import rel_l2


def test_dm_actuator_counts():
    # 21 x 21 and 81 x 81 Cartesian grids: valid if r <= D/2 + 0.7533 pitch (DeformableMirror.py:326-339)
    for nSubap, expect in ((20, 357), (80, 5209)):
        d = DMOracle(4 * nSubap, 8.0, nSubap)
        x = np.linspace(-4, 4, nSubap + 1)
        count = sum(1 for a in x for b in x if math.hypot(a, b) <= 4 + 0.7533 * 8 / nSubap)
        assert d.nValidAct == count == expect


def test_dm_modes_separable_identity():
    d = DMOracle(40, 8.0, 10, 0.35)
    modes = d.modes()
    rs = np.random.RandomState(0)
    coefs = rs.randn(d.nValidAct)
    opd = d.opd(modes, coefs)
    # independent route: Gaussian factorises on the actuator grid, OPD = Gy C Gx^T
    ux = np.unique(np.round(d.u0x, 9))
    uy = np.unique(np.round(d.u0y, 9))
    a = 1 / (2 * d.c ** 2)
    gx = np.exp(-a * (d.grid[:, None] - ux[None, :]) ** 2)
    gy = np.exp(-a * (d.grid[:, None] - uy[None, :]) ** 2)
    C = np.zeros((len(uy), len(ux)))
    C[np.searchsorted(uy, np.round(d.u0y, 9)), np.searchsorted(ux, np.round(d.u0x, 9))] = coefs
    assert rel_l2(opd, gy @ C @ gx.T) < 1e-12


def test_dm_gaussian_coupling():
    # one pitch away from an actuator the influence function equals the mechanical coupling
    d = DMOracle(40, 8.0, 10, 0.35)
    pitch_px = d.n / d.nActAlongDiameter
    assert np.isclose(np.exp(-(pitch_px ** 2) / (2 * d.c ** 2)), 0.35)


def test_warp_matches_scikit_image_golden():
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'warp_skimage.npz'))
    for k in range(int(d['n_cases'])):
        got = oat.warp_shift(d['img_%d' % k], *d['shift_%d' % k])
        assert np.max(np.abs(got - d['out_%d' % k])) < 1e-12, k


def test_integer_shift_is_exact_translation():
    img = np.random.RandomState(1).randn(10, 10)
    out = oat.warp_shift(img, 1.0, 0.0)
    assert np.array_equal(out[:, 1:], img[:, :-1])
    out = oat.warp_shift(img, 0.0, -1.0)
    assert np.array_equal(out[:-1, :], img[1:, :])


def _layer(res=12, ratio=(0.0, 1.0), seed=0):
    M = res + 2
    outer = np.ones((M, M))
    outer[1:-1, 1:-1] = 0
    inner = np.ones((M, M)) - outer
    inner[3:-3, 3:-3] = 0
    nI, nO = int(inner.sum()), int(outer.sum())
    rs = np.random.RandomState(seed)
    ms = rs.randn(M, M)
    A = rs.randn(nO, nI) * 0.01
    B = np.eye(nO) * 0.1
    return oat.LayerOracle(ms, A, B, inner, outer, RandomState(5), ratio), ms, A, B, inner, outer


def test_layer_whole_pixel_step_shifts_and_extrudes():
    L, ms, A, B, inner, outer = _layer(ratio=(0.0, 1.0))
    L.update()
    # interior moved down by one row; the new outer ring is A z + B xi with xi from the layer RNG
    assert np.array_equal(L.mapShift[2:-1, 1:-1], ms[1:-2, 1:-1])
    shifted = oat.warp_shift(ms, 0.0, 1.0)[1:-1, 1:-1]
    xi = RandomState(5).normal(size=B.shape[1])
    X = A @ shifted[inner[1:-1, 1:-1] != 0] + B @ xi
    assert np.allclose(L.mapShift[outer != 0], X)


def test_layer_subpixel_accumulates_buffer():
    L, ms, *_ = _layer(ratio=(0.4, 0.0))
    L.update()
    L.update()
    assert np.allclose(L.buff, [0.8, 0.0])
    L.update()                                  # 1.2 px: one extrusion, 0.2 px left
    assert np.allclose(L.buff, [0.2, 0.0])
    assert np.array_equal(L.mapShift[1:-1, 2:-1], ms[1:-1, 1:-2])


def test_phase_screen_statistics():
    # Ensemble structure function of 40 seeded screens vs the exact expectation of the
    # method (phaseStats.py:153-271): sum over the FFT grid and over the subharmonics the
    # reference actually adds (2 x 2 of each 3 x 3 grid, phaseStats.py:261-263) of
    # 2 PSD(f) df^2 (1 - cos(2 pi f_x r)). Pins the draw normalisation, the Re-part
    # convention and the subharmonic selection; a dropped factor 2 or a full 3 x 3
    # subharmonic set would fail.
    from oopao_amd.phaseStats import ft_sh_phase_screen
    r0, L0, N, dx = 0.15, 1e4, 128, 0.05
    Ds = []
    for seed in range(40):
        ph = ft_sh_phase_screen(r0, L0, N, dx, seed=seed)
        Ds.append(np.mean((ph[:, 8:] - ph[:, :-8]) ** 2))
    r = 8 * dx

    def psd(f):
        return 0.023 * r0 ** (-5 / 3) / ((f ** 2 + 1 / L0 ** 2) ** (11 / 6))
    df = 1 / (N * dx)
    fx = np.arange(-N / 2, N / 2) * df
    FX, FY = np.meshgrid(fx, fx)
    F = np.hypot(FX, FY)
    P = psd(np.where(F == 0, 1, F))
    P[N // 2, N // 2] = 0
    expect = np.sum(2 * P * df ** 2 * (1 - np.cos(2 * np.pi * FX * r)))
    for p in range(1, 4):
        d = 1 / (3 ** p * N * dx)
        for a, b in ((-1, -1), (0, -1), (-1, 0)):
            expect += 2 * psd(np.hypot(a * d, b * d)) * d ** 2 * (1 - np.cos(2 * np.pi * a * d * r))
    assert abs(np.mean(Ds) / expect - 1) < 0.05
    # the method under-represents the largest scales: below Kolmogorov 6.88 (r/r0)^(5/3)
    assert np.mean(Ds) < 6.88 * (r / r0) ** (5 / 3)


def test_phase_screen_seeded_reproducible():
    from oopao_amd.phaseStats import ft_sh_phase_screen
    a = ft_sh_phase_screen(0.15, 25, 36, 0.25, seed=10)
    b = ft_sh_phase_screen(0.15, 25, 36, 0.25, seed=10)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, ft_sh_phase_screen(0.15, 25, 36, 0.25, seed=11))


def test_covariance_continuity_at_zero_lag():
    # cst u^(5/6) K_{5/6}(u) -> variance as u -> 0 (phaseStats.py:39-58)
    from oopao_amd.phaseStats import makeCovarianceMatrix
    r0, L0 = 0.15, 25.0
    c = makeCovarianceMatrix(np.array([0j]), np.array([0j, 1e-7 + 0j]), L0, r0)
    assert np.isclose(c[0, 1] / c[0, 0], 1.0, rtol=1e-5)
    # von Karman phase variance closed form 0.0863 (L0/r0)^(5/3) (phaseStats.variance)
    assert np.isclose(c[0, 0], 0.0863 * (L0 / r0) ** (5 / 3), rtol=1e-3)


@pytest.mark.parametrize('M', [10, 11, 34, 326])
def test_centred_dft_matrix(M):
    # W C W^T = fftshift(fft2(fftshift(C))) (phaseStats.py:135-150) for even and odd supports
    from oopao_amd.phaseStats import centred_dft_matrix
    rs = np.random.RandomState(M)
    C = rs.randn(M, M) + 1j * rs.randn(M, M)
    W = centred_dft_matrix(M)
    ref = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(C)))
    assert np.max(np.abs(W @ C @ W.T - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize('M', [34, 35])
def test_boiling_screen_pieces(M):
    # the pieces the device boiling screen is built from reproduce the host generators exactly:
    # FFT-grid amplitude x the seed's draws -> ft_phase_screen; the 12 staged subharmonic terms from the
    # first 54 draws -> ft_sh - ft; and the oracle's restatement equals the product's host code
    from oopao_amd.phaseStats import (centred_dft_matrix, ft_amplitude, subharmonic_terms, ft_phase_screen,
                                      ft_sh_phase_screen)
    from oracle import atmosphere as oat
    r0, L0, pix, seed = 0.15, 25.0, 0.1, 100017
    g = np.random.RandomState(seed).normal(size=2 * M * M)
    cn = (g[:M * M] + 1j * g[M * M:]).reshape(M, M) * ft_amplitude(r0, L0, M, pix)
    W = centred_dft_matrix(M)
    hi = (W @ cn @ W.T).real
    ft = ft_phase_screen(r0, L0, M, pix, seed=seed)
    assert np.max(np.abs(hi - ft)) <= 1e-12 * np.max(np.abs(ft))
    sub, xc = subharmonic_terms(r0, L0, M, pix, g[:54])
    x, y = np.meshgrid(xc, xc)
    lo = sum(c[0] * np.cos(2 * np.pi * (c[2] * x + c[3] * y)) - c[1] * np.sin(2 * np.pi * (c[2] * x + c[3] * y))
             for c in sub)
    sh = ft_sh_phase_screen(r0, L0, M, pix, seed=seed)
    assert np.max(np.abs((lo - lo.mean()) + ft - sh)) <= 1e-12 * np.max(np.abs(sh))
    assert np.allclose(oat.ft_sh_phase_screen(r0, L0, M, pix, seed), sh, rtol=0, atol=1e-12 * np.max(np.abs(sh)))
    assert np.allclose(oat.ft_phase_screen(r0, L0, M, pix, seed), ft, rtol=0, atol=1e-12 * np.max(np.abs(ft)))


def test_boiling_oracle_ar1_statistics():
    # Atmosphere.apply_boiling (Atmosphere.py:387-408) is the AR(1) process m_{k+1} = a m_k + b phi_k with
    # a^2 + b^2 = 1 and independent screens phi_k: the support keeps the screens' variance and successive
    # supports have correlation a. Estimated over 400 steps of a windless layer (piston removed).
    from oracle import atmosphere as oat
    M, pix, r0, L0, dt, tb = 34, 0.25, 0.15, 25.0, 1e-3, 2e-3
    a = np.exp(-dt / tb)
    m0 = oat.ft_sh_phase_screen(r0, L0, M, pix, seed=5)
    lay = oat.LayerOracle(m0, np.zeros((1, 1)), np.zeros((1, 1)), np.zeros((M, M)), np.zeros((M, M)),
                          np.random.RandomState(0), (0.0, 0.0))
    lay.set_boiling(tb, dt, r0, L0, pix, 100000, mode=2)
    num = den = v0 = v1 = 0.0
    prev = lay.mapShift - lay.mapShift.mean()
    for k in range(400):
        lay.update()
        cur = lay.mapShift - lay.mapShift.mean()
        num += np.sum(cur * prev)
        den += np.sum(prev * prev)
        v1 += np.sum(cur * cur)
        prev = cur
    assert abs(num / den - a) < 0.05
    assert abs(v1 / den - 1) < 0.15
    assert lay.boiling_seed == 100400
    assert np.array_equal(lay.OPD, lay.mapShift[1:-1, 1:-1])


def test_dm_oracle_rows_equal_modes():
    od = DMOracle(24, 8, 6, 0.35, central_obstruction=0.3)
    rows = np.array([0, 5, 100, 287, 24 * 24 - 1])
    assert np.array_equal(od.modes_rows(rows), od.modes()[rows])
