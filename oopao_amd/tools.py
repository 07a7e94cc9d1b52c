"""Error type and helpers shared by the drop-in classes (tools/tools.py:464-483, 558-570)."""
import numpy as np


class OopaoError(Exception):
    """Exception raised for errors related to OOPAO objects (tools/tools.py:562-570)."""

    def __init__(self, string):
        self.string = string
        self.message = string
        super().__init__(self.message)


class WorkspaceAllocationError(OopaoError):
    """A device workspace could not be allocated (out of device memory, or above the byte cap set in the
    environment variable OOPAO_AMD_MAX_WORKSPACE_BYTES, which forces the failure from the host)."""


def warning(string):
    """ANSI warning print (tools/tools.py:558-559)."""
    print('\033[01m\033[33m' + 'OOPAO Warning: \n' + string + '\033[0m')


def set_binning(array, binning_factor, mode='sum'):
    """Sum/mean binning of a 2-D or 3-D numpy array (tools/tools.py:464-483)."""
    if array.shape[0] % binning_factor != 0:
        raise ValueError('Binning factor %d not compatible with the array size' % binning_factor)
    if array.ndim == 2:
        new = [int(np.round(array.shape[0] / binning_factor)), int(np.round(array.shape[1] / binning_factor))]
        shape = (new[0], array.shape[0] // new[0], new[1], array.shape[1] // new[1])
        r = array.reshape(shape)
        return r.sum(-1).sum(1) if mode == 'sum' else r.mean(-1).mean(1)
    new = [int(np.round(array.shape[0] / binning_factor)), int(np.round(array.shape[1] / binning_factor)),
           array.shape[2]]
    shape = (new[0], array.shape[0] // new[0], new[1], array.shape[1] // new[1], new[2])
    r = array.reshape(shape)
    return r.sum(-2).sum(1) if mode == 'sum' else r.mean(-2).mean(1)


def strehlMeter(PSF, tel, PSF_ref=None, zeroPaddingFactor=2, display=False, title=''):
    """Strehl ratio [%] from the OTF of a focal-plane image (tools/tools.py:102-157), on the GPU.
    Without PSF_ref the diffraction-limited image is computed after resetting the OPD, as the
    reference does (tel.resetOPD(); tel.computePSF(zeroPaddingFactor))."""
    import torch
    from .science import strehl_dev
    if PSF_ref is None:
        tel.resetOPD()
        tel.computePSF(zeroPaddingFactor)
        ref = tel._PSF
    else:
        ref = torch.as_tensor(np.asarray(PSF_ref) if not isinstance(PSF_ref, torch.Tensor) else PSF_ref,
                              device=tel.device)
    psf = PSF if isinstance(PSF, torch.Tensor) else torch.as_tensor(np.asarray(PSF), device=tel.device)
    sr = strehl_dev(psf.to(tel.device), ref.to(tel.device), tel.resolution, zeroPaddingFactor)
    if display:
        print('Strehl ratio [%] : ', sr)
    return sr


# ------------------------------------------------------------------------------------------------
# Scintillation diagnostics (host numpy, analysis of n x n maps; tools/tools.py:574-696)
# ------------------------------------------------------------------------------------------------
def radial_profile(data, max_radius):
    """Azimuthal mean of a 2-D map about its centre (shape/2), integer radius bins (tools.py:574-585)."""
    y, x = np.indices(data.shape)
    cy, cx = data.shape[0] / 2, data.shape[1] / 2
    r = np.sqrt((x - cx) ** 2 + (y - cy) ** 2).astype(int).ravel()
    sums = np.bincount(r, data.ravel())
    counts = np.bincount(r)
    return (sums / np.maximum(counts, 1))[:max_radius]


def compute_rytov_variance(scintillation, pupil):
    """4 var(chi), chi = log(I / mean I) / 2 over the pupil pixels with I > 1e-10 (tools.py:587-599)."""
    if scintillation is None:
        return 0.0
    I = np.asarray(scintillation)[np.asarray(pupil) > 0]
    I = I[I > 1e-10]
    chi = 0.5 * np.log(I / np.mean(I))
    return 4.0 * np.var(chi)


def compute_theoretical_pupil_log_amp_variance(wavelength, D_tel, L_source, L0, Cn2_input):
    """Pupil-averaged log-amplitude variance (tools.py:602-622): 5.20 R^(5/3) k0^2 times the double
    integral over altitude z (0..L_source) and normalised frequency k (0..200) of
    Cn2(z) k^(-14/3) J1(k)^2 sin^2(z k^2 / (2 k0 R^2)) (1 + (2 pi R / (L0 k))^2)^(-11/6)."""
    from scipy.integrate import quad
    from scipy.special import j1
    k0 = 2 * np.pi / wavelength
    R = D_tel / 2.0

    def f_k(k, z):
        if k < 1e-9:
            return 0.0
        return (k ** (-14 / 3) * j1(k) ** 2 * np.sin(z * k ** 2 / (2 * k0 * R ** 2)) ** 2
                * (1 + (2 * np.pi * R / (L0 * k)) ** 2) ** (-11 / 6))

    def f_z(z):
        c = Cn2_input(z) if callable(Cn2_input) else Cn2_input
        return c * quad(f_k, 0, 200, args=(z,), limit=100)[0]

    return 5.20 * R ** (5 / 3) * k0 ** 2 * quad(f_z, 0, L_source)[0]


def compute_theoretical_rytov(wavelength, r0, altitudes, fractionalR0):
    """Layered Rytov variance 2.25 k0^(7/6) Cn2_int sum_i w_i z_i^(5/6) with w = fractionalR0^(5/3)
    normalised and Cn2_int = r0^(-5/3) / (0.423 k0^2) (tools.py:624-650)."""
    k0 = 2 * np.pi / wavelength
    w = np.asarray(fractionalR0, dtype=float) ** (5 / 3)
    w = w / np.sum(w)
    z = np.asarray(altitudes, dtype=float)
    cn2_int = r0 ** (-5 / 3) / (0.423 * k0 ** 2)
    return float(np.sum(w * z ** (5 / 6)) * cn2_int * 2.25 * k0 ** (7 / 6))


def compute_dsp(phase, scintillation, resolution, pixel_size):
    """Radial phase and log-amplitude power spectra with a 2-D Tukey(0.2) window (tools.py:653-685):
    chi = log(I / mean_centre I + 1e-12) / 2, phase minus its centre mean, window-weighted mean removed,
    |fft2|^2 pixel_size^2 / sum(window^2), azimuthal means; returns (kappa, dsp_phi, dsp_chi)."""
    from scipy.signal.windows import tukey
    phase = np.asarray(phase, dtype=float)
    scint = np.asarray(scintillation, dtype=float)
    N = phase.shape[0]
    w1 = tukey(N, alpha=0.2)
    w = np.outer(w1, w1)
    c = slice(N // 4, 3 * N // 4)
    chi = 0.5 * np.log(scint / np.mean(scint[c, c]) + 1e-12)
    phi = phase - np.mean(phase[c, c])
    out = []
    for a in (phi, chi):
        aw = a * w
        aw = aw - np.sum(aw) / np.sum(w) * w
        out.append(radial_profile(np.fft.fftshift(np.abs(np.fft.fft2(aw)) ** 2) * pixel_size ** 2 / np.sum(w ** 2),
                                  N // 2))
    kappa = np.fft.fftshift(np.fft.fftfreq(N, pixel_size))[N // 2:]
    return kappa, out[0], out[1]


def compute_theoretical_psd(kappa, wavelength, r0, L0, altitude):
    """Von Karman phase PSD 0.023 r0^(-5/3) (kappa^2 + L0^-2)^(-11/6) split by the Fresnel filter
    cos^2 / sin^2(altitude (2 pi kappa)^2 / (2 k)) into phase and log-amplitude parts (tools.py:687-696)."""
    k = 2 * np.pi / wavelength
    kappa = np.asarray(kappa, dtype=float)
    psd = 0.023 * r0 ** (-5 / 3) * (kappa ** 2 + (1 / L0) ** 2) ** (-11 / 6)
    arg = altitude * (2 * np.pi * kappa) ** 2 / (2 * k)
    return psd * np.cos(arg) ** 2, psd * np.sin(arg) ** 2
