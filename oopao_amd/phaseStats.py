"""Von Karman phase screens and covariances (phaseStats.py:19-271), host-side init code.

Phase screens are drawn from numpy's legacy RandomState exactly as the
reference does (same seeds, same draw order), so a seeded screen is the
reference's screen up to FFT rounding. Used at initialisation only; the
per-frame update runs on the GPU (oopao_amd/csrc/atm.hip).
"""
import math

import numpy as np
import scipy.special
from numpy.random import RandomState


def gamma(x):
    return math.gamma(x)


def covariance_constants(L0, r0):
    """(cst, var) of the von Karman phase covariance (phaseStats.py:39-58, 72-80)."""
    L0r0 = (L0 / r0) ** (5. / 3)
    cst = (24. * gamma(6. / 5) / 5) ** (5. / 6) * (gamma(11. / 6) / ((2. ** (5. / 6)) * np.pi ** (8. / 3))) * L0r0
    var = (24. * gamma(6. / 5) / 5) ** (5. / 6) * (gamma(11. / 6) * gamma(5. / 6) / (2 * np.pi ** (8. / 3))) * L0r0
    return cst, var


def makeCovarianceMatrix(rho1, rho2, L0, r0_def):
    """Phase covariance between complex coordinate sets rho1, rho2 (phaseStats.py:72-132)."""
    rho = np.abs(rho1[:, None] - rho2[None, :])
    cst, var = covariance_constants(L0, r0_def)
    out = np.full(rho.shape, var)
    index = np.where(rho != 0)
    u = 2 * np.pi * rho[index] / L0
    out[index] = cst * u ** (5. / 6) * scipy.special.kv(5. / 6, u)
    return out


def extrusion_rings(resolution, n_extra_pixel, D):
    """Ring geometry of the Assemat extrusion for a layer of ``resolution`` px and diameter D
    (Atmosphere.py:335-349): on the (resolution + n_extra_pixel)^2 support, the outer ring (the
    border pixels that an extrusion writes) and the inner ring (the n_extra_pixel-wide band inside it
    that conditions the draw), as masks and as complex coordinates u + i v on a pitch of
    D / (resolution - 1). Returns (outer, inner, innerZ, outerZ)."""
    M = resolution + n_extra_pixel
    outer = np.ones((M, M))
    outer[1:-1, 1:-1] = 0
    inner = np.ones((M, M)) - outer
    inner[1 + n_extra_pixel:-1 - n_extra_pixel, 1 + n_extra_pixel:-1 - n_extra_pixel] = 0
    x = np.linspace(0, resolution + 1, resolution + 2) * D / (resolution - 1)
    u, v = np.meshgrid(x, x)
    return outer, inner, u[inner != 0] + 1j * v[inner != 0], u[outer != 0] + 1j * v[outer != 0]


def extrusion_operators(ZZt, ZXt, XXt, ZZt_inv, r0_def, r0):
    """Extrusion operators at Fried parameter r0 from the covariances computed at r0_def
    (Atmosphere.py:358-363, 1027-1035): with c = (r0_def / r0)^(5/3), A = (c ZXt)^T (ZZt_inv / c)
    (independent of r0) and B = chol(c XXt - A c ZXt), so that an extruded ring X = A Z + B xi has
    the von Karman covariance c XXt jointly with the inner ring Z. Returns (A, B)."""
    c = (r0_def / r0) ** (5 / 3)
    A = np.matmul((ZXt * c).T, ZZt_inv / c)
    B = np.linalg.cholesky(XXt * c - np.matmul(A, ZXt * c))
    return A, B


def centred_dft_matrix(M):
    """W [M, M] complex128 with W C W^T = fftshift(fft2(fftshift(C))) (the ift2 of phaseStats.py:135-150):
    W = P F P for the DFT F and the fftshift permutation P (shift by h = M // 2), i.e.
    W[k, l] = exp(-2 pi i ((k - h) mod M)((l + h) mod M) / M), the exponent reduced exactly mod M."""
    h = M // 2
    k = np.arange(M)
    r = np.outer((k - h) % M, (k + h) % M) % M
    return np.exp(-2j * np.pi * r / M)


def ft_amplitude(r0, L0, N, delta, l0=1e-10):
    """sqrt(PSD) df of ft_phase_screen's FFT grid (phaseStats.py:176-195), PSD[N/2, N/2] = 0: the
    screen's Fourier coefficients are (g_re + i g_im) times this."""
    delta = float(delta)
    del_f = 1. / (N * delta)
    fx = np.arange(-N / 2., N / 2.) * del_f
    fx, fy = np.meshgrid(fx, fx)
    f = np.sqrt(fx ** 2 + fy ** 2)
    fm = 5.92 / l0 / (2 * np.pi)
    psd = 0.023 * r0 ** (-5. / 3.) * np.exp(-1 * ((f / fm) ** 2)) / (((f ** 2) + (1. / L0) ** 2) ** (11. / 6))
    psd[int(N / 2), int(N / 2)] = 0
    return np.sqrt(psd) * del_f


def subharmonic_terms(r0, L0, resolution, pixel_size, draws, l0=1e-10):
    """The 12 subharmonic terms of ft_sh_phase_screen (phaseStats.py:236-266) as rows (Re c, Im c, fx, fy)
    of c exp(2 pi i (fx x + fy y)), and the grid coordinates x = y; ``draws`` are the first 54 normal draws
    of RandomState(seed) (per level: 9 real then 9 imaginary parts, row-major 3 x 3)."""
    pixel_size = float(pixel_size)
    D = resolution * pixel_size
    g = np.asarray(draws, np.float64)[:54].reshape(3, 2, 3, 3)
    rows = []
    for p in range(1, 4):
        del_f = 1 / (3 ** p * D)
        fxv = np.arange(-1, 2) * del_f
        fx, fy = np.meshgrid(fxv, fxv)
        f = np.sqrt(fx ** 2 + fy ** 2)
        fm = 5.92 / l0 / (2 * np.pi)
        psd = 0.023 * r0 ** (-5. / 3) * np.exp(-1 * (f / fm) ** 2) / ((f ** 2 + (1. / L0) ** 2) ** (11. / 6))
        psd[1, 1] = 0
        cn = (g[p - 1, 0] + 1j * g[p - 1, 1]) * np.sqrt(psd) * del_f
        for i in range(0, 2):
            for j in range(0, 2):
                rows.append((cn[i, j].real, cn[i, j].imag, fx[i, j], fy[i, j]))
    coords = np.arange(-resolution / 2, resolution / 2) * pixel_size
    return np.asarray(rows, np.float64), coords


def subharmonic_table(r0, L0, resolution, pixel_size, l0=1e-10):
    """The 12 subharmonic terms of ft_sh_phase_screen as rows (re index, im index, a, fx, fy): the term's
    coefficient is (g[re] + i g[im]) a for the screen's normal draws g, whose first 54 are the
    subharmonics' (per level 9 real then 9 imaginary parts, row-major 3 x 3; phaseStats.py:236-266)."""
    unit = np.zeros(54)
    rows = []
    for p in range(3):
        for i in range(2):
            for j in range(2):
                re, im = 18 * p + 3 * i + j, 18 * p + 9 + 3 * i + j
                unit[:] = 0
                unit[re] = 1.0
                t = subharmonic_terms(r0, L0, resolution, pixel_size, unit)[0][4 * p + 2 * i + j]
                rows.append((re, im, t[0], t[2], t[3]))
    return np.asarray(rows, np.float64)


def ift2(G):
    """fftshift(fft2(fftshift(G))) (phaseStats.py:135-150)."""
    return np.fft.fftshift(np.fft.fft2(np.fft.fftshift(G)))


def ft_phase_screen(r0, L0, N, delta, l0=1e-10, seed=None):
    """FFT von Karman screen in radians at 500 nm (phaseStats.py:153-198)."""
    delta = float(delta)
    if seed is None:
        import random
        import time
        seed = int(random.SystemRandom(time.time()).random() * 100000)
    rs = RandomState(seed)
    del_f = 1. / (N * delta)
    fx = np.arange(-N / 2., N / 2.) * del_f
    fx, fy = np.meshgrid(fx, fx)
    f = np.sqrt(fx ** 2 + fy ** 2)
    fm = 5.92 / l0 / (2 * np.pi)
    f0 = 1. / L0
    psd = 0.023 * r0 ** (-5. / 3.) * np.exp(-1 * ((f / fm) ** 2)) / (((f ** 2) + (f0 ** 2)) ** (11. / 6))
    psd[int(N / 2), int(N / 2)] = 0
    cn = (rs.normal(size=(N, N)) + 1j * rs.normal(size=(N, N))) * np.sqrt(psd) * del_f
    return ift2(cn).real


def ft_sh_phase_screen(r0, L0, resolution, pixel_size, l0=1e-10, seed=None):
    """FFT screen plus 3 levels of 2 x 2 subharmonics (phaseStats.py:201-271)."""
    pixel_size = float(pixel_size)
    if seed is None:
        import random
        import time
        seed = int(random.SystemRandom(time.time()).random() * 100000)
    rs = RandomState(seed)
    D = resolution * pixel_size
    phs_hi = ft_phase_screen(r0, L0, resolution, pixel_size, seed=seed)
    coords = np.arange(-resolution / 2, resolution / 2) * pixel_size
    x, y = np.meshgrid(coords, coords)
    phs_lo = np.zeros(phs_hi.shape)
    for p in range(1, 4):
        del_f = 1 / (3 ** p * D)
        fx = np.arange(-1, 2) * del_f
        fx, fy = np.meshgrid(fx, fx)
        f = np.sqrt(fx ** 2 + fy ** 2)
        fm = 5.92 / l0 / (2 * np.pi)
        f0 = 1. / L0
        psd = 0.023 * r0 ** (-5. / 3) * np.exp(-1 * (f / fm) ** 2) / ((f ** 2 + f0 ** 2) ** (11. / 6))
        psd[1, 1] = 0
        cn = (rs.normal(size=(3, 3)) + 1j * rs.normal(size=(3, 3))) * np.sqrt(psd) * del_f
        SH = np.zeros((resolution, resolution), dtype='complex')
        for i in range(0, 2):
            for j in range(0, 2):
                SH += cn[i, j] * np.exp(1j * 2 * np.pi * (fx[i, j] * x + fy[i, j] * y))
        phs_lo = phs_lo + SH
    phs_lo = phs_lo.real - phs_lo.real.mean()
    return phs_lo + phs_hi
