"""Host scintillation diagnostics (oopao_amd.tools; reference tools/tools.py:574-696, Source.py:460-521),
pinned by closed forms and identities rather than by a restatement."""
import numpy as np

from oopao_amd import tools


def test_radial_profile_constant_and_ring():
    assert np.allclose(tools.radial_profile(np.full((16, 16), 3.0), 8), 3.0)
    y, x = np.indices((32, 32))
    ring = (np.sqrt((x - 16) ** 2 + (y - 16) ** 2).astype(int) == 5).astype(float)
    prof = tools.radial_profile(ring, 16)
    assert prof[5] == 1.0 and np.count_nonzero(prof) == 1


def test_rytov_variance_is_four_var_chi():
    rs = np.random.RandomState(0)
    chi = 0.1 * rs.randn(40, 40)
    pupil = np.ones((40, 40))
    pupil[:5] = 0
    got = tools.compute_rytov_variance(np.exp(2 * chi), pupil)
    assert abs(got - 4 * np.var(chi[5:])) < 1e-12
    assert tools.compute_rytov_variance(None, pupil) == 0.0


def test_theoretical_rytov_closed_form():
    lam, r0 = 500e-9, 0.15
    k = 2 * np.pi / lam
    one = tools.compute_theoretical_rytov(lam, r0, [8000.0], [1.0])
    ref = 2.25 * k ** (7 / 6) * r0 ** (-5 / 3) / (0.423 * k ** 2) * 8000.0 ** (5 / 6)
    assert abs(one / ref - 1) < 1e-12
    # equal fractions: weights 1/2; a ground layer adds nothing
    two = tools.compute_theoretical_rytov(lam, r0, [0.0, 8000.0], [0.5, 0.5])
    assert abs(two / (0.5 * ref) - 1) < 1e-12


def test_theoretical_psd_split():
    kappa = np.linspace(0.01, 5, 50)
    phi0, chi0 = tools.compute_theoretical_psd(kappa, 500e-9, 0.15, 25, 0.0)
    vk = 0.023 * 0.15 ** (-5 / 3) * (kappa ** 2 + 1 / 25 ** 2) ** (-11 / 6)
    assert np.allclose(phi0, vk, rtol=1e-14) and np.all(chi0 == 0)
    phi, chi = tools.compute_theoretical_psd(kappa, 500e-9, 0.15, 25, 10000.0)
    assert np.allclose(phi + chi, vk, rtol=1e-12)


def test_dsp_peaks_at_the_phase_frequency():
    N, dx, k0 = 64, 0.1, 6
    x = np.arange(N) * dx
    phase = np.cos(2 * np.pi * k0 / (N * dx) * x)[None, :] * np.ones((N, 1))
    kappa, dphi, dchi = tools.compute_dsp(phase, np.ones((N, N)), N, dx)
    assert len(kappa) == N // 2 and abs(kappa[k0] - k0 / (N * dx)) < 1e-12
    assert int(np.argmax(dphi)) == k0
    assert np.max(np.abs(dchi)) < 1e-20


def test_pupil_log_amp_variance_linear_in_cn2():
    a = tools.compute_theoretical_pupil_log_amp_variance(500e-9, 8.0, 10000.0, 25.0, 1e-16)
    b = tools.compute_theoretical_pupil_log_amp_variance(500e-9, 8.0, 10000.0, 25.0, 2e-16)
    assert a > 0 and abs(b / a - 2) < 1e-9
    assert tools.compute_theoretical_pupil_log_amp_variance(500e-9, 8.0, 0.0, 25.0, 1e-16) == 0.0
