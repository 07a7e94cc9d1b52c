"""numpy restatement of OOPAO's KL modal basis (CPU oracle, test infrastructure only).

Follows calibration/compute_KL_modal_basis.py:14-332 with the helpers of
calibration/ao_cockpit_psim.py (VK_DSP_up 155-192, DO_HHt 847-1141,
build_SpecificBasis_C 582-621, build_SeedBasis_C 659-689, build_KLBasis
724-745) and the Zernike specific modes (Zernike.py:17-66), along the
reference's own arithmetic route: full complex FFTs of the IFs centred in the
SZ x SZ grid, conj(F)^T diag(PSD df^2) F, and numpy SVDs.

Pins (tests/test_oracle_kl.py): the PSD constant against its closed form; HHt
against the spatial-domain double sum sum_x sum_y IF_i(x) IF_j(y) C(x - y) with
C the inverse transform of PSD df^2; the invariants the reference prints
(orthonormality SpM^T DELTA SpM / tpup = I, SB^T DELTA SB / tpup = I,
KL^T DELTA KL / tpup = I, diagonal KL covariance); Zernike tip/tilt as exact
linear ramps over the pupil.
"""
import math

import numpy as np


def zernike_ptt(pupil, D, nModes=3):
    """Noll modes 1..nModes (piston, tip, tilt, ...) as [n, n, nModes] (Zernike.py:17-66)."""
    n = pupil.shape[0]
    rows, cols = np.where(pupil > 0)
    off = (n + n % 2 - 1) / 2
    X = (rows - off) / n * D
    Y = (cols - off) / n * D
    R = np.sqrt(X ** 2 + Y ** 2)
    R = R / R.max()
    th = np.arctan2(Y, X)
    out = np.zeros((n, n, nModes))
    for j in range(1, nModes + 1):
        nn = int((-1. + np.sqrt(8 * (j - 1) + 1)) / 2.)
        p = j - nn * (nn + 1) / 2.
        k = nn % 2
        m = int((p + k) / 2.) * 2 - k
        if m != 0 and j % 2 != 0:
            m = -m
        am = abs(m)
        Rr = np.zeros_like(R)
        for i in range((nn - am) // 2 + 1):
            Rr += ((-1) ** i * math.factorial(nn - i) /
                   (math.factorial(i) * math.factorial((nn + am) // 2 - i) * math.factorial((nn - am) // 2 - i))
                   ) * R ** (nn - 2 * i)
        if m == 0:
            Z = np.sqrt(nn + 1) * Rr
        elif m > 0:
            Z = np.sqrt(2 * (nn + 1)) * Rr * np.cos(am * th)
        else:
            Z = np.sqrt(2 * (nn + 1)) * Rr * np.sin(am * th)
        if nn != 0:
            Z = (Z - Z.mean()) / np.std(Z - Z.mean())
        out[rows, cols, j - 1] = Z
    return out


VK_CST = (math.gamma(11. / 6.) ** 2 / (2. * np.pi ** (11. / 3.))) * (24. * math.gamma(6. / 5.) / 5.) ** (5. / 6.)


def vk_psd(diameter, r0, L0, size, dim, pupil):
    """Piston-filtered von Karman PSD, centred (ao_cockpit_psim.py:155-192, Pcor = 1)."""
    dom = pupil.shape[0]
    P = np.zeros([size, size])
    P[size // 2 - dom // 2:size // 2 + dom // 2, size // 2 - dom // 2:size // 2 + dom // 2] = pupil
    fmax = dim / diameter * 0.5
    L0 = 1.e6 if L0 == 0 else L0
    f = (np.arange(size) - size / 2) / (size / 2) * fmax
    fr2 = f[None, :] ** 2 + f[:, None] ** 2
    FT = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(P)))
    pterm = np.abs(FT) ** 2
    psd = VK_CST * r0 ** (-5. / 3.) * (fr2 + L0 ** -2) ** (-11. / 6.) * (1. - pterm / pterm.max())
    return psd, fmax / (size / 2)


def hht(IF, psd, df, SZ):
    """HHt[i, j] = Re sum_f conj(F_i) F_j PSD df^2, F = fft2 of the IF centred in SZ x SZ
    (ao_cockpit_psim.py:847-1141)."""
    nact, n, _ = IF.shape
    A = np.zeros((nact, SZ, SZ))
    A[:, SZ // 2 - n // 2:SZ // 2 + n // 2, SZ // 2 - n // 2:SZ // 2 + n // 2] = IF
    F = np.fft.fft2(A).reshape(nact, SZ * SZ)
    d = (np.fft.fftshift(psd) * df ** 2).reshape(-1)
    return np.real(np.conj(F) @ (F * d[None, :]).T)


def _nmax(s, lim):
    if lim > 1:
        return int(lim)
    return int(np.max(np.where(s / np.max(s) > lim)))


def specific_basis(Tspm, IFma, DELTA, lim, ortho=True):
    """ao_cockpit_psim.py:582-621."""
    tpup = IFma.shape[0]
    nact = IFma.shape[1]
    ud, sd, vdh = np.linalg.svd(DELTA)
    nmax = _nmax(sd, lim)
    MEV = np.zeros((nact, nact))
    MEV[:nmax, :nmax] = np.diag(1. / sd[:nmax])
    iDELTA = vdh.T @ MEV @ ud.T
    cmd = iDELTA @ IFma.T @ Tspm
    if not ortho:
        return cmd
    _, R = np.linalg.qr(IFma @ cmd)
    return cmd @ np.linalg.inv(R) * np.sqrt(tpup)


def seed_basis(SpM, DELTA, lim, tpup, return_spectrum=False):
    """ao_cockpit_psim.py:659-689. The truncation keeps the first nmax eigenvectors, nmax being the
    last index above lim: when eigenvalues nmax - 1 and nmax are equal (a degenerate pair split by
    the cut) the kept vector is an arbitrary member of that pair, in the reference as here."""
    nact = DELTA.shape[0]
    nspm = SpM.shape[1]
    G = np.eye(nact) - SpM @ np.linalg.inv(SpM.T @ DELTA @ SpM) @ SpM.T @ DELTA
    Mp, D2, _ = np.linalg.svd(G.T @ DELTA @ G)
    nmax = _nmax(D2, lim) if lim != 0 else nact - nspm
    M = Mp[:, :nmax] @ np.diag(np.sqrt(1. / D2[:nmax])) * np.sqrt(tpup)
    if return_spectrum:
        return G @ M, D2
    return G @ M


def kl_basis(HHt, SB, DELTA, nmoKL):
    """ao_cockpit_psim.py:724-745."""
    SB = SB[:, :nmoKL]
    DSB = SB.T @ DELTA @ SB
    iD = np.linalg.inv(DSB)
    Cp = iD @ SB.T @ HHt @ SB @ iD.T
    _, Sc, VcT = np.linalg.svd(Cp)
    return SB @ VcT.T, Sc


def compute_kl(IF2D, pupil, D, r0, L0, lim=1e-3, remove_piston=True, SZ=None, nZer=3, nmo=None):
    """Full pipeline (compute_KL_modal_basis.py:38-332, minimF False, ortho_spm True).
    IF2D [nAct, n, n]. Returns dict(BASIS, SpM, SB, KL, Sc, HHt, DELTA, tpup)."""
    n = pupil.shape[0]
    IF = IF2D * pupil[None]
    nact = IF.shape[0]
    SZ = 2 * n if SZ is None else SZ
    psd, df = vk_psd(D, r0, L0, SZ, n, pupil)
    H = hht(IF, psd, df, SZ)
    idx = np.where(pupil == 1)
    tpup = len(idx[0])
    IFma = IF[:, idx[0], idx[1]].T
    Tspm = zernike_ptt(pupil, D, nZer)[idx[0], idx[1], :]
    DELTA = IFma.T @ IFma
    SpM = specific_basis(Tspm, IFma, DELTA, lim)
    SB, D2 = seed_basis(SpM, DELTA, lim, tpup, return_spectrum=True)
    nmo = nact if nmo is None else nmo
    nmoKL = min(nmo, SB.shape[1])
    KL, Sc = kl_basis(H, SB, DELTA, nmoKL)
    BASIS = np.concatenate([SpM, KL], axis=1)
    if remove_piston:
        BASIS = BASIS[:, 1:]
    return dict(BASIS=BASIS, SpM=SpM, SB=SB, KL=KL, Sc=Sc, HHt=H, DELTA=DELTA, tpup=tpup, psd=psd, df=df,
                IFma=IFma, D2=D2)
