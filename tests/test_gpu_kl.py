"""GPU: the float64 KL basis on the device against the numpy oracle (oracle/kl.py).

Reference: calibration/compute_KL_modal_basis.py:14-332 and ao_cockpit_psim.py:155-192, 582-745,
847-1141. Eigenvectors are defined up to sign (and up to a rotation inside a degenerate eigenspace) in
every LAPACK implementation, so modes are compared through their Gram overlaps: within each group of
equal variances (relative gap < 1e-7) the overlap matrix must be orthogonal to 1e-6. Variances,
HHt and DELTA agree with the oracle to 1e-9 relative; orthonormality to 1e-8. The seed-basis cut
keeps the first nmax eigenvectors; when it splits a degenerate pair the kept member is arbitrary (in
the reference too), so mode-by-mode parity is checked at a threshold near the default that splits no
pair, and the default threshold is checked through the invariants.
"""
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import kl as okl
from oracle.dm import DMOracle
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def _objects(n, nSubap):
    from oopao_amd import Telescope, DeformableMirror
    tel = Telescope(n, 8, 1e-3, centralObstruction=0)
    dm = DeformableMirror(tel, nSubap, 0.35, print_dm_properties=False)
    atm = SimpleNamespace(r0=0.15, L0=25.0)
    return tel, dm, atm


def _oracle_lim(n, nSubap, lim):
    d = DMOracle(n, 8.0, nSubap, 0.35)
    from oracle.pwfs import telescope_pupil
    pup = telescope_pupil(n).astype(float)
    IF2D = np.moveaxis(d.modes().reshape(n, n, -1), -1, 0)
    return okl.compute_kl(IF2D, pup, 8.0, 0.15, 25.0, lim=lim)


def _oracle(n, nSubap):
    return _oracle_lim(n, nSubap, 1e-3)


def _groups(s, tol=1e-7):
    groups, cur = [], [0]
    for i in range(1, len(s)):
        if abs(s[i] - s[i - 1]) <= tol * s[0]:
            cur.append(i)
        else:
            groups.append(cur)
            cur = [i]
    groups.append(cur)
    return groups


def _safe_lim(D2, lim, tol=1e-6):
    """A truncation threshold near ``lim`` whose cut does not split a degenerate pair of the seed
    spectrum (the kept vector of a split pair is an arbitrary member of it, in the reference too)."""
    s = D2 / D2.max()
    j = int(np.max(np.where(s > lim)))
    while j > 2 and (s[j - 1] - s[j] < tol * s[0] or s[j] - s[j + 1] < tol * s[0]):
        j -= 1
    return 0.5 * (s[j] + s[j + 1])


def _device_basis(tel, atm, dm, n, lim):
    from oopao_amd.calibration.compute_KL_modal_basis import compute_M2C
    return compute_M2C(tel, atm, dm, remove_piston=True, SZ=2 * n, nZer=3, NDIVL=1, recompute_cov=True,
                       save_output=False, lim_inversion=lim, display=False)


@pytest.mark.parametrize('n,nSubap', [(24, 6), (48, 12), (80, 20)])
def test_kl_basis_matches_oracle(n, nSubap):
    import torch
    from oopao_amd.calibration.compute_KL_modal_basis import hht_dev, vk_psd
    tel, dm, atm = _objects(n, nSubap)
    r0 = _oracle(n, nSubap)
    # influence functions and HHt through the rfft / real-GEMM route
    IF = np.moveaxis(dm.modes.reshape(n, n, -1), -1, 0)
    pup = np.asarray(tel.pupil, float)
    IFd = dm.influence_functions_f64().T.reshape(-1, n, n)
    assert rel_l2(IFd.cpu().numpy(), IF) < 1e-6   # float32 resident modes vs float64
    psd, df, _ = vk_psd(8.0, 0.15, 25.0, 2 * n, n, pup)
    H = hht_dev(IFd * torch.from_numpy(pup).cuda()[None], psd, df, 2 * n, block=16).cpu().numpy()
    err = rel_l2(H, r0['HHt'])
    print('HHt rel_l2', err)
    assert err < 1e-9
    # default threshold: a valid KL basis (orthonormal, diagonal covariance) of a seed space that
    # differs from the oracle's at most by the arbitrary member of a pair split by the cut
    B = _device_basis(tel, atm, dm, n, 1e-3)
    tp, D, Ho = r0['tpup'], r0['DELTA'], r0['HHt']
    assert B.shape == r0['BASIS'].shape
    err = np.max(np.abs(B.T @ D @ B / tp - np.eye(B.shape[1])))
    print('orthonormality error', err)
    assert err < 1e-8
    KL = B[:, 2:]
    cov = KL.T @ Ho @ KL / tp ** 2
    off = np.max(np.abs(cov - np.diag(np.diag(cov)))) / np.max(np.diag(cov))
    print('KL covariance off-diagonal', off)
    assert off < 1e-9
    sv = np.linalg.svd(r0['KL'].T @ D @ KL / tp, compute_uv=False)
    assert np.sum(np.abs(sv - 1) > 1e-6) <= 1
    # a threshold that splits no pair: the same basis as the oracle, mode by mode up to sign /
    # rotation inside equal-variance groups, same variances
    lim = _safe_lim(r0['D2'], 1e-3)
    r = _oracle_lim(n, nSubap, lim)
    B = _device_basis(tel, atm, dm, n, lim)
    assert B.shape == r['BASIS'].shape
    for k in range(2):
        a, b = B[:, k], r['BASIS'][:, k]
        assert min(rel_l2(a, b), rel_l2(a, -b)) < 1e-8
    KL = B[:, 2:]
    Sc_dev = np.diag(KL.T @ Ho @ KL) / tp ** 2
    err = np.max(np.abs(Sc_dev - r['Sc']) / r['Sc'].max())
    print('KL variance error', err, 'lim', lim)
    assert err < 1e-9
    O = r['KL'].T @ D @ KL / tp
    worst = 0.0
    for g in _groups(r['Sc']):
        sv = np.linalg.svd(O[np.ix_(g, g)], compute_uv=False)
        worst = max(worst, np.max(np.abs(sv - 1)))
    print('eigenspace overlap error', worst, 'modes', KL.shape[1])
    assert worst < 1e-6
