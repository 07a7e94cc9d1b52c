"""Shared builders for tests: seeded inputs and oracle systems."""
import numpy as np

from oracle import pwfs as opw


def rel_l2(a, b):
    """||a - b|| / ||b|| in float64 (complex128 when either input is complex)."""
    a, b = np.asarray(a), np.asarray(b)
    dt = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a = a.astype(dt)
    b = b.astype(dt)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def small_system(n=16, nSubap=4, modulation=2, T=16, precision='fp64', lightRatio=0.1):
    """Tiny PWFS (N = 64) with 4 px per subaperture, separation nSubap, edge nSubap/2."""
    pup = opw.telescope_pupil(n)
    flux = opw.flux_map(pup, opw.n_photon('I3', 8), 1e-3, 8.0)
    o = opw.PyramidOracle(n, nSubap, modulation, lightRatio, pup, flux, 0.8e-6,
                          n_pix_separation=nSubap, n_pix_edge=nSubap // 2,
                          nTheta_user_defined=T, precision=precision)
    return o


def config1_system(precision='fp64', n_jobs=1):
    """Config 1: 80 px pupil, 20 x 20 subapertures, 48 modulation points at 3 lambda/D, I3 band."""
    n = 80
    pup = opw.telescope_pupil(n)
    flux = opw.flux_map(pup, opw.n_photon('I3', 8), 1e-3, 8.0)
    return opw.PyramidOracle(n, 20, 3, 0.5, pup, flux, 0.8e-6, n_pix_separation=20, n_pix_edge=10,
                             nTheta_user_defined=48, precision=precision, n_jobs=n_jobs)


def config2_system(precision='fp64', n_jobs=16):
    """Config 2: 320 px pupil, 80 x 80 subapertures, 72 points at 3 lambda/D (N = 1280)."""
    n = 320
    pup = opw.telescope_pupil(n)
    flux = opw.flux_map(pup, opw.n_photon('I3', 8), 1e-3, 8.0)
    return opw.PyramidOracle(n, 80, 3, 0.5, pup, flux, 0.8e-6, n_pix_separation=80, n_pix_edge=40,
                             nTheta_user_defined=72, precision=precision, n_jobs=n_jobs)


def random_phase(n, seed, rms=1.0, pupil=None):
    rs = np.random.RandomState(seed)
    ph = rs.randn(n, n)
    # smooth a little so the phase looks like a wavefront, not white noise
    k = np.fft.fftfreq(n)
    filt = np.exp(-(k[:, None] ** 2 + k[None, :] ** 2) * 40)
    ph = np.real(np.fft.ifft2(np.fft.fft2(ph) * filt))
    ph *= rms / ph.std()
    if pupil is not None:
        ph = ph * pupil
    return ph
