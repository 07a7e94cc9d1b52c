"""CPU: pins of the Bi-O Edge oracle (oracle/bioedge.py)."""
import numpy as np

from oracle import pwfs as opw
from oracle.bioedge import BioEdgeGeometry, BioEdgeOracle, bioedge_masks


def test_masks_split_energy_and_half_planes():
    g = BioEdgeGeometry(32, 8, 4)
    for gw, gl in ((0, False), (3, False), (2, 3)):
        M = bioedge_masks(g, grey_width=gw, grey_length=gl)
        a = np.abs(M) ** 2
        assert np.allclose(a[0] + a[1], 1, atol=1e-14)
        assert np.allclose(a[2] + a[3], 1, atol=1e-14)
    M = bioedge_masks(g)
    N = g.N
    assert np.all(np.abs(M[0])[:, :N // 2] == 1) and np.all(np.abs(M[0])[:, N // 2:] == 0)
    assert np.all(np.abs(M[2])[:N // 2, :] == 1) and np.all(np.abs(M[2])[N // 2:, :] == 0)


def test_raw_flux_conservation():
    n = 32
    pup = opw.telescope_pupil(n)
    flux = opw.flux_map(pup, opw.n_photon('I3', 8), 1e-3, 8.0)
    o = BioEdgeOracle(n, 8, 2, 0.1, pup, flux, n_pix_separation=4, nTheta_user_defined=8, grey_width=2)
    ph = np.random.RandomState(0).randn(n, n) * 0.3 * pup
    assert np.isclose(o.raw(ph).sum(), flux.sum(), rtol=1e-12)


def test_mask_shift_rolls_image():
    n = 32
    pup = opw.telescope_pupil(n)
    flux = opw.flux_map(pup, opw.n_photon('I3', 8), 1e-3, 8.0)
    kw = dict(n_pix_separation=4, nTheta_user_defined=8)
    o0 = BioEdgeOracle(n, 8, 0, 0.0, pup, flux, **kw)
    o1 = BioEdgeOracle(n, 8, 0, 0.0, pup, flux, sx=[0, 0, 0, 2], **kw)     # mask A (index 0) takes sx[3]
    ph = np.random.RandomState(1).randn(n, n) * 0.2 * pup
    r0, r1 = o0.raw(ph), o1.raw(ph)
    N = o0.g.N
    shift = 2 * n // 8                        # sx * norma raw pixels, norma = n / nSubap
    t0, t1 = r0[:N, :N], r1[:N, :N]
    assert np.allclose(t1, np.roll(t0, shift, axis=1), rtol=1e-10, atol=1e-10 * t0.max())
    assert np.allclose(r0[N:, :], r1[N:, :], rtol=1e-12, atol=1e-12 * t0.max())
