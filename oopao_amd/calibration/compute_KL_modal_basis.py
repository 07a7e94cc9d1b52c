"""KL modal basis of a DM for von Karman turbulence, in float64 on the GPU.

Follows calibration/compute_KL_modal_basis.py:14-332 and the helpers of
calibration/ao_cockpit_psim.py it calls (VK_DSP_up 155-192, DO_HHt 847-1141,
build_SpecificBasis_C 582-621, build_SeedBasis_C 659-689, build_KLBasis
724-745), with Zernike piston/tip/tilt as specific modes (Zernike.py:17-66):

* HHt[i, j] = Re sum_f conj(F_i) F_j PSD(f) df^2, F_i = FFT_SZ of the i-th
  pupil-masked influence function. The IFs are real, so the sum is taken over
  the half spectrum of a real-to-complex transform (rocFFT) with weights 1 on
  the self-conjugate columns kx = 0, SZ/2 and 2 elsewhere:
  HHt = G G^T with G = [Re F, Im F] sqrt(w), one float64 GEMM (rocBLAS).
* DELTA = IF^T IF over the pupil pixels (float64 GEMM).
* The reference's SVDs of symmetric positive semi-definite matrices (DELTA,
  G^T DELTA G, the KL covariance Cp) are eigen-decompositions; they run as
  float64 symmetric eigensolvers on the device (rocSOLVER), sorted by
  decreasing eigenvalue like the SVD. The truncation rule keeps the
  reference's index arithmetic: nmax = the last index with s / max(s) > lim,
  and the first nmax values are kept.

Eigenvectors carry an arbitrary sign (and, for equal eigenvalues, an arbitrary
rotation inside their eigenspace) in any LAPACK implementation, the
reference's included; the basis is defined by its spectrum, its orthonormality
(KL^T DELTA KL / tpup = I) and its diagonal covariance.
"""
import math

import numpy as np
import torch

from ..tools import OopaoError
from ..Zernike import Zernike


def vk_psd(diameter, r0, L0, size, dim, pupil):
    """Piston-filtered von Karman phase PSD on a size x size grid, centred, and its df
    (ao_cockpit_psim.py:155-192 with Pcor = 1)."""
    dom = pupil.shape[0]
    pad = np.zeros([size, size])
    pad[size // 2 - dom // 2:size // 2 + dom // 2, size // 2 - dom // 2:size // 2 + dom // 2] = pupil
    fmax = dim / diameter * 0.5
    if L0 == 0:
        L0 = 1.e6
    f = np.linspace(-size / 2, size / 2 - 1, size) / (size / 2) * fmax
    fx, fy = np.meshgrid(f, f)
    fr2 = fx ** 2 + fy ** 2
    cst = (math.gamma(11. / 6.) ** 2 / (2. * np.pi ** (11. / 3.))) * (24. * math.gamma(6. / 5.) / 5.) ** (5. / 6.)
    shifted = np.roll(np.roll(pad, size // 2, 0), size // 2, 1)
    pterm = np.abs(np.roll(np.roll(np.fft.fft2(shifted), size // 2, 0), size // 2, 1)) ** 2
    pterm0 = pterm / np.max(pterm)
    psd = cst * r0 ** (-5 / 3) * (fr2 + (1. / L0) ** 2) ** (-11 / 6) * (1. - pterm0)
    df = abs(f[size // 2] - f[size // 2 - 1])
    return psd, df, pterm


def hht_dev(IF, psd, df, SZ, block=256):
    """Covariance of the IFs' projections on the turbulence, HHt [nAct, nAct] float64 on the device.
    IF: [nAct, n, n] float64 device tensor (pupil applied); psd: centred [SZ, SZ] (host)."""
    nact, n, _ = IF.shape
    if n > SZ:
        raise OopaoError('FFT size SZ=%d smaller than the IF size %d' % (SZ, n))
    dev = IF.device
    half = SZ // 2 + 1
    w = np.fft.fftshift(psd)[:, :half] * df ** 2
    c = np.full(half, 2.0)
    c[0] = 1.0
    if SZ % 2 == 0:
        c[-1] = 1.0
    sw = torch.from_numpy(np.sqrt(np.maximum(w * c[None, :], 0.0))).to(dev)
    L = SZ * half
    G = torch.empty((nact, 2 * L), dtype=torch.float64, device=dev)
    buf = torch.zeros((min(block, nact), SZ, SZ), dtype=torch.float64, device=dev)
    for k0 in range(0, nact, block):
        k1 = min(nact, k0 + block)
        b = buf[:k1 - k0]
        b.zero_()
        b[:, :n, :n] = IF[k0:k1]
        F = torch.fft.rfft2(b) * sw
        G[k0:k1, :L] = F.real.reshape(k1 - k0, L)
        G[k0:k1, L:] = F.imag.reshape(k1 - k0, L)
        del F
    del buf
    H = G @ G.T
    del G
    torch.cuda.empty_cache()
    return 0.5 * (H + H.T)


def _eig_desc(A):
    """Eigen-decomposition of a symmetric matrix, eigenvalues in decreasing order."""
    s, V = torch.linalg.eigh(0.5 * (A + A.T))
    return s.flip(0), V.flip(1)


def _nmax(s, lim):
    if lim > 1:
        return int(lim)
    r = (s / s.max()).cpu().numpy()
    return int(np.max(np.where(r > lim)))


def specific_basis(Tspm, IFma, DELTA, lim, ortho=True):
    """Commands of the specific modes (ao_cockpit_psim.py:582-621, no regularisation)."""
    tpup = IFma.shape[0]
    s, U = _eig_desc(DELTA)
    nmax = _nmax(s, lim)
    Uk = U[:, :nmax]
    iDELTA = Uk @ torch.diag(1.0 / s[:nmax]) @ Uk.T
    cmd = iDELTA @ (IFma.T @ Tspm)
    if not ortho:
        return cmd
    _, R = torch.linalg.qr(IFma @ cmd)
    return torch.linalg.solve_triangular(R, cmd, upper=True, left=False) * math.sqrt(tpup)


def seed_basis(SpM, DELTA, lim, tpup):
    """DM eigenmodes orthogonal to the specific modes (ao_cockpit_psim.py:659-689)."""
    nact = DELTA.shape[0]
    nspm = SpM.shape[1]
    eye = torch.eye(nact, dtype=torch.float64, device=DELTA.device)
    A = SpM.T @ DELTA @ SpM
    G = eye - SpM @ torch.linalg.solve(A, SpM.T @ DELTA)
    D2, Mp = _eig_desc(G.T @ DELTA @ G)
    nmax = _nmax(D2, lim) if lim != 0 else nact - nspm
    M = Mp[:, :nmax] * torch.sqrt(1.0 / D2[:nmax])[None, :] * math.sqrt(tpup)
    return G @ M


def _spd_inverse(A):
    """Inverse of a symmetric positive definite matrix through its eigen-decomposition (GEMMs and
    the symmetric eigensolver only; no large triangular solves)."""
    s, V = torch.linalg.eigh(0.5 * (A + A.T))
    return (V / s[None, :]) @ V.T


def kl_basis(HHt, SB, DELTA, nmoKL):
    """Diagonalise the turbulence covariance in the seed basis (ao_cockpit_psim.py:724-745).
    Returns (KL commands, variances Sc)."""
    SB = SB[:, :nmoKL]
    DSB = SB.T @ DELTA @ SB
    iD = _spd_inverse(DSB)
    Cp = iD @ (SB.T @ HHt @ SB) @ iD.T
    Sc, Vc = _eig_desc(Cp)
    return SB @ Vc, Sc


def influence_functions_2d(dm, pupil_dev):
    """IFs [nAct, n, n] float64 on the device, pupil applied (compute_KL_modal_basis.py:74-79)."""
    n = dm.resolution
    modes = dm.influence_functions_f64()          # [n*n, nAct]
    IF = modes.T.reshape(-1, n, n) * pupil_dev[None]
    return IF


def compute_M2C(telescope, atmosphere, deformableMirror, param=None, nameFolder=None, nameFile=None,
                remove_piston=False, HHtName=None, baseName=None, SpM_2D=None, nZer=3, SZ=None,
                mem_available=None, NDIVL=None, computeSpM=True, ortho_spm=True, computeSB=True,
                computeKL=True, minimF=False, P2F=None, alpha=None, beta=None, lim_SpM=None, lim_SB=None,
                nmo=None, IF_2D=None, IFma=None, returnSB=False, returnHHt_PSD_df=False,
                recompute_cov=False, extra_name='', save_output=True, lim_inversion=1e-3, display=True):
    """Modal basis (DM commands per mode) [nAct, nModes] (compute_KL_modal_basis.py:38-332)."""
    if minimF:
        raise OopaoError('force-minimised bases (minimF=True, stiffness matrix from FITS) are not supported')
    if save_output:
        if display:
            print('Saving the basis to FITS is not supported; the basis is returned only.')
    dev = deformableMirror.device
    n = telescope.resolution
    pupil = np.asarray(telescope.pupil, dtype=np.float64)
    pupil_dev = torch.from_numpy(pupil).to(dev)
    if nmo is None:
        nmo = deformableMirror.nValidAct
    if IF_2D is None:
        IF = influence_functions_2d(deformableMirror, pupil_dev)
    else:
        IF = torch.as_tensor(np.asarray(IF_2D, np.float64), device=dev) * pupil_dev[None]
    nact = IF.shape[0]
    if display:
        print('Computing Specific Modes ...')
    if nZer is not None and SpM_2D is None:
        Z = Zernike(telescope, J=nZer)
        Z.computeZernike(telescope, remove_piston=0)
        SpM_2D = Z.modesFullRes
    nspm = SpM_2D.shape[2]
    if SZ is None:
        SZ = int(2 * n)
    if display:
        print('COMPUTING VON KARMAN 2D PSD...')
    psd, df, _ = vk_psd(telescope.D, atmosphere.r0, atmosphere.L0, SZ, n, pupil)
    if display:
        print('COMPUTING COV MAT HHt...')
    block = max(1, nact // NDIVL) if NDIVL else 256
    HHt = hht_dev(IF, psd, df, SZ, block=min(block, 512))
    idx = torch.from_numpy(np.flatnonzero(pupil.reshape(-1) == 1)).to(dev)
    tpup = int(idx.numel())
    IFma = IF.reshape(nact, n * n)[:, idx].T.contiguous()           # [tpup, nAct]
    Tspm = torch.from_numpy(np.ascontiguousarray(np.moveaxis(SpM_2D, 2, 0).reshape(nspm, n * n))).to(dev)
    Tspm = Tspm[:, idx].T.contiguous()
    DELTA = IFma.T @ IFma
    lim = lim_inversion if lim_SpM is None else lim_SpM
    SpM = specific_basis(Tspm, IFma, DELTA, lim, ortho_spm)
    if display:
        err = torch.max(torch.abs(SpM.T @ DELTA @ SpM / tpup - torch.eye(nspm, dtype=torch.float64, device=dev)))
        print('Orthonormality error for SpM = ', float(err))
    SB = seed_basis(SpM, DELTA, lim_inversion, tpup)
    nSB = SB.shape[1]
    if not computeKL:
        BASIS = torch.cat([SpM, SB], dim=1)
        if remove_piston:
            BASIS = BASIS[:, 1:]
        out = BASIS.cpu().numpy()
        return (out, SB.cpu().numpy()) if returnSB else out
    if nmo > nSB:
        print('WARNING: Number of modes requested too high, taking the maximum value possible!')
        nmoKL = nSB
    else:
        nmoKL = nmo
    KL, Sc = kl_basis(HHt, SB, DELTA, nmoKL)
    if display:
        DKL = KL.T @ DELTA @ KL
        err = torch.max(torch.abs(DKL / tpup - torch.eye(nmoKL, dtype=torch.float64, device=dev)))
        print('Orthonormality error for ' + str(nmoKL) + ' modes of the KL Basis = ', float(err))
    BASIS = torch.cat([SpM, KL], dim=1)
    if remove_piston:
        BASIS = BASIS[:, 1:]
        if display:
            print('Piston removed from the modal basis!')
    out = BASIS.cpu().numpy()
    if returnHHt_PSD_df:
        return out, HHt.cpu().numpy(), psd, df
    if returnSB:
        return out, SB.cpu().numpy()
    return out


def compute_KL_basis(tel, atm, dm, lim=1e-3, remove_piston=True, n_batch=1):
    """KL basis of the DM (compute_KL_modal_basis.py:14-35): [nAct, nAct - 1] with piston removed."""
    return compute_M2C(telescope=tel, atmosphere=atm, deformableMirror=dm, param=None, nameFolder=None,
                       nameFile=None, remove_piston=remove_piston, HHtName=None, baseName=None,
                       mem_available=None, minimF=False, nmo=None, ortho_spm=True, SZ=int(2 * tel.resolution),
                       nZer=3, NDIVL=n_batch, recompute_cov=True, save_output=False, lim_inversion=lim,
                       display=False)
