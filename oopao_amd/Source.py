"""Source: carrier of the optical path difference, phase and flux (Source.py:12-549).

Same constructor, tags, operators and properties as the reference. The OPD,
OPD_no_pupil, mask, fluxMap, scintillation and intensity maps are kept on the
GPU (float32, batch-major for cubes); reading one of the public attributes
returns a float64 numpy array in the reference's layout ([n,n] or [n,n,B]).
"""
import numpy as np
import torch

from ._device import default_device, to_dev, to_host
from .tools import OopaoError

_CST = 368
# Source.py:342-424 -- [wavelength (m), bandwidth (m), zero point (photons/m^2/s)]
_PHOT = {
    'U': [0.360e-6, 0.070e-6, 1.96e12 / _CST], 'B': [0.440e-6, 0.100e-6, 5.38e12 / _CST],
    'V0': [0.500e-6, 0.090e-6, 3.64e12 / _CST], 'V': [0.550e-6, 0.090e-6, 3.31e12 / _CST],
    'R': [0.640e-6, 0.150e-6, 4.01e12 / _CST], 'R2': [0.650e-6, 0.300e-6, 7.9e12 / _CST],
    'R3': [0.600e-6, 0.300e-6, 8.56e12 / _CST], 'R4': [0.670e-6, 0.300e-6, 7.66e12 / _CST],
    'I': [0.790e-6, 0.150e-6, 2.69e12 / _CST], 'I1': [0.700e-6, 0.033e-6, 0.67e12 / _CST],
    'I2': [0.750e-6, 0.033e-6, 0.62e12 / _CST], 'I3': [0.800e-6, 0.033e-6, 0.58e12 / _CST],
    'I4': [0.700e-6, 0.100e-6, 2.02e12 / _CST], 'I5': [0.850e-6, 0.100e-6, 1.67e12 / _CST],
    'I6': [1.000e-6, 0.100e-6, 1.42e12 / _CST], 'I7': [0.850e-6, 0.300e-6, 5.00e12 / _CST],
    'I8': [0.750e-6, 0.100e-6, 1.89e12 / _CST], 'I9': [0.850e-6, 0.300e-6, 5.00e12 / _CST],
    'I10': [0.900e-6, 0.300e-6, 4.72e12 / _CST], 'J': [1.215e-6, 0.260e-6, 1.90e12 / _CST],
    'J2': [1.550e-6, 0.260e-6, 1.49e12 / _CST], 'H': [1.654e-6, 0.290e-6, 1.05e12 / _CST],
    'Kp': [2.1245e-6, 0.351e-6, 0.62e12 / _CST], 'Ks': [2.157e-6, 0.320e-6, 0.55e12 / _CST],
    'K': [2.179e-6, 0.410e-6, 0.70e12 / _CST], 'K0': [2.000e-6, 0.410e-6, 0.76e12 / _CST],
    'K1': [2.400e-6, 0.410e-6, 0.64e12 / _CST], 'L': [3.547e-6, 0.570e-6, 2.5e11 / _CST],
    'M': [4.769e-6, 0.450e-6, 8.4e10 / _CST], 'Na': [0.589e-6, 0, 3.3e12 / _CST],
    'EOS': [1.064e-6, 0, 3.3e12 / _CST], 'IR1310': [1.310e-6, 0, 2e12 / _CST],
}


class Source:
    def __init__(self, optBand: str, magnitude: float, coordinates: list = [0, 0], altitude: float = np.inf,
                 laser_coordinates: list = [0, 0], Na_profile=None, fwhm_spot_up=None,
                 display_properties: bool = True, chromatic_shift: list = None, device=None):
        self.device = default_device() if device is None else torch.device(device)
        self.is_initialized = False
        self.display_properties = display_properties
        self.__updating_flux = False
        self.optBand = optBand
        self.wavelength, self.bandwidth, self.zeroPoint = self.photometry(optBand)
        self._magnitude = magnitude
        self.mask = 1
        self._nPhoton = self.zeroPoint * 10 ** (-0.4 * magnitude)
        self.tag = 'source'
        self.altitude = altitude
        self.coordinates = coordinates
        self.laser_coordinates = laser_coordinates
        self.chromatic_shift = chromatic_shift
        if Na_profile is not None and fwhm_spot_up is not None:
            self.Na_profile = Na_profile
            self.fwhm_spot_up = fwhm_spot_up
            self.altitude = np.sum(Na_profile[0, :] * Na_profile[1, :])
            self.type = 'LGS'
        else:
            self.type = 'NGS'
        self.optical_path = [[self.type + '(' + self.optBand + ')', self]]
        self.through_atm = False
        self.inAsterism = False
        self.ast_idx = -1
        self._OPD = None
        self._OPD_no_pupil = None
        self._fluxMap = None
        self._scintillation = None
        self._scintillation_no_pupil = None
        self._intensity = None
        self.phase_filtered = None
        self.amplitude_filtered = None
        self.var_chi = None
        self.var_rytov = None
        self.tel = None
        if self.display_properties:
            print(self)
        self.is_initialized = True

    # ---------------------------------------------------------------- operators
    def __pow__(self, obj):
        """Re-propagation (Source.py:167-174)."""
        obj.src = self
        self.optical_path = [[self.type + '(' + self.optBand + ')', self]]
        self.reset()
        self.through_atm = False
        self * obj
        return self

    def __mul__(self, obj):
        """Propagation (Source.py:176-184)."""
        if obj is not None:
            if isinstance(obj, list):
                for o in obj:
                    self * o
            else:
                obj.relay(self)
        return self

    def reset(self):
        self.mask = 1
        self._OPD = None
        self._OPD_no_pupil = None
        self.phase_filtered = None
        self.amplitude_filtered = None
        self._scintillation = None
        self._scintillation_no_pupil = None
        self._refresh_intensity()
        self.var_chi = None
        self.var_rytov = None

    def print_optical_path(self):
        if self.optical_path is not None:
            print(' ~~> '.join(p[0] for p in self.optical_path))

    # ------------------------------------------------------------ device state
    @property
    def mask(self):
        return self._mask

    @mask.setter
    def mask(self, val):
        self._mask = val
        self._mask_t = None

    def _set_mask(self, host, dev):
        """Pupil mask from a Telescope: host array plus its device mirror (no upload)."""
        self._mask = host
        self._mask_t = dev

    def _mask_dev(self):
        if np.isscalar(self._mask):
            return None
        if self._mask_t is None:
            self._mask_t = torch.from_numpy(np.asarray(self._mask, np.float32)).to(self.device)
        return self._mask_t

    def _set_OPD_no_pupil_dev(self, t):
        """Set OPD_no_pupil (device tensor [n,n] or [B,n,n]) and OPD = OPD_no_pupil * mask."""
        self._OPD_no_pupil = t
        if t is None:
            self._OPD = None
            return
        m = self._mask_dev()
        self._OPD = t if m is None else t * m

    def _refresh_intensity(self):
        """intensity = fluxMap * scintillation (Source.py:206-228)."""
        flux, scint = self._fluxMap, self._scintillation
        if flux is None:
            self._intensity = None
        elif scint is None:
            self._intensity = flux
        else:
            self._intensity = flux * scint

    # ------------------------------------------------------- public properties
    @property
    def OPD(self):
        return to_host(self._OPD)

    @OPD.setter
    def OPD(self, val):
        if val is None:
            self._OPD = None
            self._OPD_no_pupil = None
        else:
            self._OPD = to_dev(val, self.device)

    @property
    def OPD_no_pupil(self):
        return to_host(self._OPD_no_pupil)

    @OPD_no_pupil.setter
    def OPD_no_pupil(self, val):
        if val is not None:
            self._set_OPD_no_pupil_dev(to_dev(val, self.device))
        else:
            self._OPD_no_pupil = None

    @property
    def phase(self):
        """OPD * 2 pi / lambda (Source.py:294-296)."""
        return None if self._OPD is None else to_host(self._OPD) * 2 * np.pi / self.wavelength

    @phase.setter
    def phase(self, val):
        self._OPD = to_dev(np.asarray(val) * self.wavelength / (2 * np.pi), self.device) \
            if not isinstance(val, torch.Tensor) else (val * (self.wavelength / (2 * np.pi))).to(self.device, torch.float32)

    @property
    def phase_no_pupil(self):
        return None if self._OPD_no_pupil is None else to_host(self._OPD_no_pupil) * 2 * np.pi / self.wavelength

    @phase_no_pupil.setter
    def phase_no_pupil(self, val):
        self.OPD_no_pupil = np.asarray(val) * self.wavelength / (2 * np.pi)

    @property
    def fluxMap(self):
        return to_host(self._fluxMap)

    @fluxMap.setter
    def fluxMap(self, val):
        self._fluxMap = None if val is None else to_dev(val, self.device)
        self._refresh_intensity()

    @property
    def scintillation(self):
        return to_host(self._scintillation)

    @scintillation.setter
    def scintillation(self, val):
        if val is None:
            self._scintillation = None
            self._scintillation_no_pupil = None
        else:
            self._scintillation = to_dev(val, self.device)
        self._refresh_intensity()

    @property
    def scintillation_no_pupil(self):
        return to_host(self._scintillation_no_pupil)

    @scintillation_no_pupil.setter
    def scintillation_no_pupil(self, val):
        if val is None:
            self._scintillation_no_pupil = None
            return
        t = to_dev(val, self.device)
        self._scintillation_no_pupil = t
        m = self._mask_dev()
        self._scintillation = t if m is None else t * m
        self._refresh_intensity()

    @property
    def intensity(self):
        return to_host(self._intensity)

    @intensity.setter
    def intensity(self, val):
        self._intensity = None if val is None else to_dev(val, self.device)

    @property
    def var(self):
        """Phase variance inside the pupil (set by Telescope.relay, Telescope.py:221)."""
        if self._OPD is None or self.tel is None or self._OPD.dim() != 2:
            return None
        ph = self._OPD.double() * (2 * np.pi / self.wavelength)
        sel = ph[self.tel._pupil_dev > 0]
        return float(sel.var(unbiased=False))

    # ------------------------------------------------------------- photometry
    def photometry(self, arg):
        if isinstance(arg, str):
            if arg in _PHOT:
                w, b, z = _PHOT[arg]
                return w, b, z
            raise OopaoError('Wrong name for the photometry object')
        raise OopaoError('The photometry object takes a string as an input')

    @property
    def nPhoton(self):
        return self._nPhoton

    @nPhoton.setter
    def nPhoton(self, val):
        if self.__updating_flux:
            return
        self.__updating_flux = True
        self._nPhoton = val
        self._magnitude = -2.5 * np.log10(val / self.zeroPoint)
        print('Flux updated, magnitude is %2i and flux is %.2e' % (self._magnitude, self._nPhoton))
        self.__updating_flux = False

    @property
    def magnitude(self):
        return self._magnitude

    @magnitude.setter
    def magnitude(self, val):
        if self.__updating_flux:
            return
        self.__updating_flux = True
        self._magnitude = val
        self._nPhoton = self.zeroPoint * 10 ** (-0.4 * self._magnitude)
        print('Flux updated, magnitude is %2i and flux is %.2e' % (self._magnitude, self._nPhoton))
        self.__updating_flux = False

    # ------------------------------------------------------------------ print
    # ------------------------------------------------- scintillation diagnostics (host, Source.py:460-521)
    @property
    def rytov_variance(self):
        """Measured Rytov variance of the scintillation over the telescope pupil (Source.py:460-466)."""
        from .tools import compute_rytov_variance
        return compute_rytov_variance(self.scintillation, self.tel.pupil)

    def get_dsp(self, telescope):
        """Radial phase and log-amplitude power spectra of the unmasked maps (Source.py:468-482)."""
        from .tools import compute_dsp
        return compute_dsp(self.phase_no_pupil, self.scintillation_no_pupil, telescope.resolution,
                           telescope.pixelSize)

    def get_theoretical_pupil_variance(self, D_tel, L_source, L0, Cn2_input):
        from .tools import compute_theoretical_pupil_log_amp_variance
        return compute_theoretical_pupil_log_amp_variance(self.wavelength, D_tel, L_source, L0, Cn2_input)

    def get_theoretical_rytov(self, atm):
        from .tools import compute_theoretical_rytov
        return compute_theoretical_rytov(self.wavelength, atm.r0, atm.altitude, atm.fractionalR0)

    def get_theoretical_dsp(self, kappa, r0, L0, altitude):
        from .tools import compute_theoretical_psd
        return compute_theoretical_psd(kappa, self.wavelength, r0, L0, altitude)

    def print_properties(self):
        print(self)

    def properties(self) -> dict:
        self.prop = dict()
        self.prop['type'] = f"{'Source':<25s}|{self.type:^9s}"
        self.prop['wavelength'] = f"{'Wavelength [m]':<25s}|{self.wavelength:^9.1e}"
        self.prop['zenith'] = f"{'Zenith [arcsec]':<25s}|{self.coordinates[0]:^9.2f}"
        self.prop['azimuth'] = f"{'Azimuth [°]':<25s}|{self.coordinates[1]:^9.2f}"
        self.prop['altitude'] = f"{'Altitude [m]':<25s}|{self.altitude:^9.2f}"
        self.prop['magnitude'] = f"{'Magnitude':<25s}|{self._magnitude:^9.2f}"
        self.prop['flux'] = f"{'Flux [photon/m²/s]':<25s}|{self._nPhoton:^9.1e}"
        return self.prop

    def __repr__(self):
        self.properties()
        n_char = len(max(self.prop.values(), key=len))
        return (f'\n{" Source ":-^{n_char}}\n' + ''.join(v + '\n' for v in self.prop.values())
                + f'{"":-^{n_char}}\n')
