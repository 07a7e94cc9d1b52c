"""Telescope: pupil, flux and the `tel*obj`, `tel+atm` operators (Telescope.py:19-617).

The pupil is built exactly as the reference (Telescope.py:226-246) on the host,
then mirrored on the GPU; relay() sets the source's mask and flux map on the
device (Telescope.py:199-224).
"""
import copy

import numpy as np
import torch

from ._device import default_device, to_host
from .tools import OopaoError, warning


class Telescope:
    def __init__(self, resolution: int, diameter: float, samplingTime: float = 0.001,
                 centralObstruction: float = 0., fov: float = 0., pupil: bool = None,
                 pupilReflectivity: float = 1., display_optical_path: bool = False, device=None):
        self.device = default_device() if device is None else torch.device(device)
        self.precision = np.float64
        self.isInitialized = False
        self.resolution = resolution
        self.D = diameter
        self.initial_D = diameter
        self.initial_resolution = resolution
        self.pixelSize = self.D / self.resolution
        self.centralObstruction = centralObstruction
        self.fov = fov
        self.rad2arcsec = (180 / np.pi) * 3600
        self.fov_rad = fov / self.rad2arcsec
        self.samplingTime = samplingTime
        self.isPetalFree = False
        self.index_pixel_petals = None
        self.optical_path = None
        self.user_defined_pupil = pupil
        self.pupilReflectivity = np.float64(pupilReflectivity)
        self.set_pupil()
        self.src = None
        self.tag = 'telescope'
        self.isPaired = False
        self.spatialFilter = None
        self.display_optical_path = display_optical_path
        self.coronagraph_diameter = None
        print(self)
        self.isInitialized = True
        self.apply_off_axis_tip_tilt = True
        self.initial_pupil = self.pupil.copy()
        self.static_OPD = np.zeros([self.resolution, self.resolution])

    def set_pupil(self):
        """Circular pupil with central obstruction (Telescope.py:226-246)."""
        if self.user_defined_pupil is None:
            D = self.resolution + 1
            x = np.linspace(-self.resolution / 2, self.resolution / 2, self.resolution)
            xx, yy = np.meshgrid(x, x)
            circle = xx ** 2 + yy ** 2
            obs = circle >= (self.centralObstruction * D / 2) ** 2
            pupil = (circle < (D / 2) ** 2) * obs
        else:
            warning('User-defined pupil, the central obstruction will not be taken into account...')
            pupil = self.user_defined_pupil.copy()
        refl = self.pupilReflectivity
        self.pupil = pupil
        self.pupilReflectivity = (self.pupil * refl).astype(np.float64)

    def apply_spiders(self, angle, thickness_spider, offset_X=None, offset_Y=None):
        """Telescope.apply_spiders (Telescope.py:425-459): removes, from the default pupil, the strips of
        width thickness_spider [m] through the centre (shifted by offset_X/Y [m]) at each angle [deg],
        each strip kept on one side of the centre only."""
        self.isInitialized = False
        if thickness_spider > 0:
            self.set_pupil()
            pup = self.pupil.copy()
            max_offset = self.centralObstruction * self.D / 2 - thickness_spider / 2
            offset_X = np.zeros(len(angle)) if offset_X is None else np.asarray(offset_X, dtype=float)
            offset_Y = np.zeros(len(angle)) if offset_Y is None else np.asarray(offset_Y, dtype=float)
            if np.max(np.abs(offset_X)) >= max_offset or np.max(np.abs(offset_Y)) > max_offset:
                warning('The spider offsets are too large! Weird things could happen!')
            h = self.resolution // 2
            x = np.linspace(-self.D / 2, self.D / 2, self.resolution)
            for a, ox, oy in zip(angle, offset_X, offset_Y):
                a = (a + 90) % 360
                X, Y = np.meshgrid(x + ox, x)
                Y = Y + oy
                dist = np.abs(X * np.cos(np.deg2rad(a)) + Y * np.sin(np.deg2rad(-a)))
                # the strip only extends from the centre towards its own direction
                if 0 <= a < 90:
                    dist[:h, :] = thickness_spider
                elif 90 <= a < 180:
                    dist[:, :h] = thickness_spider
                elif 180 <= a < 270:
                    dist[h:, :] = thickness_spider
                else:
                    dist[:, h:] = thickness_spider
                pup = pup * (dist > thickness_spider / 2)
            self.isInitialized = True
            self.pupil = pup
        else:
            warning('Thickness is <=0, returning default pupil')
            self.set_pupil()

    @property
    def pupil(self):
        return self._pupil

    @pupil.setter
    def pupil(self, val):
        self._pupil = np.asarray(val).astype(bool)
        self.pixelArea = np.sum(self._pupil)
        self.pupilLogical = np.where(self._pupil.reshape(self.resolution ** 2) > 0)
        self.pupilReflectivity = self._pupil.astype(np.float64)
        self._pupil_dev = torch.from_numpy(self._pupil.astype(np.float32)).to(self.device)
        if self.isInitialized:
            warning('A new pupil is now considered, its reflectivity is considered to be uniform. '
                    'Assign the proper reflectivity map to tel.pupilReflectivity if required.')

    @property
    def pupilReflectivity(self):
        return self._pupilReflectivity

    @pupilReflectivity.setter
    def pupilReflectivity(self, val):
        self._pupilReflectivity = val
        if np.ndim(val) == 2:
            self._refl_dev = torch.from_numpy(np.asarray(val, np.float32)).to(self.device)
        else:
            self._refl_dev = None

    def relay(self, src):
        """Telescope.relay (Telescope.py:199-224) on device tensors."""
        self.src = src
        if src.tag == 'source':
            self.src_list = [src]
        elif src.tag == 'asterism':
            self.src_list = src.src
        for s in self.src_list:
            s.optical_path.append([self.tag, self])
            s.tel = self
            s._set_mask(self.pupil, self._pupil_dev)
            if s._OPD is None:
                s._set_OPD_no_pupil_dev(torch.zeros((self.resolution, self.resolution),
                                                    dtype=torch.float32, device=self.device))
            if s._OPD.dim() == 2:
                s._set_OPD_no_pupil_dev(s._OPD_no_pupil if s._OPD_no_pupil is not None else s._OPD)
            else:
                s._set_OPD_no_pupil_dev(torch.zeros((self.resolution, self.resolution),
                                                    dtype=torch.float32, device=self.device))
            if s._scintillation is None or s._scintillation_no_pupil is None:
                s._scintillation_no_pupil = torch.ones((self.resolution, self.resolution),
                                                       dtype=torch.float32, device=self.device)
            s._scintillation = s._scintillation_no_pupil * self._pupil_dev
            flux = self.src_flux_scale(s)
            s._fluxMap = self._refl_dev * flux
            s._refresh_intensity()

    # --------------------------------------------------------------- science
    def computePSF(self, zeroPaddingFactor=2, detector=None, img_resolution=None):
        """PSF of the coupled on-axis source (Telescope.py:251-361), computed on the GPU
        (oopao_amd.science.psf_dev). Sets PSF, PSF_norma and the image axes."""
        from .science import psf_dev, psf_geometry
        if detector is not None:
            zeroPaddingFactor = detector.psf_sampling
            if detector.resolution is not None:
                img_resolution = detector.resolution
            else:
                img_resolution = int(zeroPaddingFactor * self.resolution)
                detector.resolution = img_resolution
        if img_resolution is None:
            img_resolution = int(zeroPaddingFactor * self.resolution)
        if self.src is None:
            raise OopaoError('The telescope was not coupled to any source object! Make sure to couple it with an '
                             'src object using src*tel')
        if self.src.tag != 'source':
            raise OopaoError('PSFs of asterisms are not computed by oopao_amd')
        src = self.src
        if np.any(np.asarray(src.coordinates)[:1] != 0) and self.apply_off_axis_tip_tilt:
            raise OopaoError('off-axis sources are not supported by oopao_amd')
        if src._OPD is None or src._OPD.dim() != 2:
            raise OopaoError('computePSF needs a single 2-D OPD on the source (propagate it through the telescope)')
        amp = torch.sqrt(src._intensity) * self._pupil_dev
        if self._refl_dev is not None:
            amp = amp * self._refl_dev
        psf_geometry(self.resolution, zeroPaddingFactor, img_resolution)     # validates the sizes
        self._PSF = psf_dev(amp.float(), src._OPD.float(), 2 * np.pi / src.wavelength, self.resolution,
                            zeroPaddingFactor, img_resolution, self.device)
        lam_d = src.wavelength / self.D
        n_pix = int(img_resolution)
        self.xPSF_rad = [-lam_d * (n_pix / 2 / zeroPaddingFactor), lam_d * (n_pix / 2 / zeroPaddingFactor)]
        self.yPSF_rad = list(self.xPSF_rad)
        self.xPSF_arcsec = [v * self.rad2arcsec for v in self.xPSF_rad]
        self.yPSF_arcsec = list(self.xPSF_arcsec)
        self.PSF_list = self._PSF

    def PropagateField(self, amplitude, phase, zeroPaddingFactor, img_resolution=None):
        """Telescope.PropagateField (Telescope.py:362-423) without coronagraph: PSF = the centred
        img_resolution window of |fft2(pad(amplitude exp(i phase)) phasor) / N|^2, binned by the
        integer oversampling used when zeroPaddingFactor < 2 (returned). Computed on the GPU."""
        from .science import psf_dev, psf_geometry
        if self.coronagraph_diameter is not None:
            raise OopaoError('coronagraphic PSFs are not computed by oopao_amd')
        n = self.pupil.shape[0]
        over = psf_geometry(n, zeroPaddingFactor, img_resolution)[0]
        amp = torch.as_tensor(np.asarray(amplitude, dtype=np.float32), device=self.device)
        ph = torch.as_tensor(np.asarray(phase, dtype=np.float32), device=self.device)
        if amp.dim() == 0:
            amp = amp.expand(n, n)
        if tuple(amp.shape) != (n, n) or tuple(ph.shape) != (n, n):
            raise OopaoError('amplitude and phase must be %d x %d arrays' % (n, n))
        self._PSF = psf_dev(amp.contiguous(), ph.contiguous(), 1.0, n, zeroPaddingFactor, img_resolution,
                            self.device)
        return over

    @property
    def PSF(self):
        p = getattr(self, '_PSF', None)
        return None if p is None else to_host(p)

    @PSF.setter
    def PSF(self, val):
        self._PSF = None if val is None else torch.as_tensor(np.asarray(val), dtype=torch.float32,
                                                              device=self.device)

    @property
    def PSF_norma(self):
        p = getattr(self, '_PSF', None)
        return None if p is None else to_host(p / p.max())

    def src_flux_scale(self, src):
        """nPhoton * samplingTime * pixel area (Telescope.py:223)."""
        return float(src.nPhoton * self.samplingTime * (self.D / self.resolution) ** 2)

    # ------------------------------------------------------------------ OPD
    @property
    def OPD(self):
        op = self.src.OPD
        if op is not None and op.ndim == 2:
            self.mean_removed_OPD = (op - np.mean(op[np.where(self.pupil == 1)])) * self.pupil
        return op

    @OPD.setter
    def OPD(self, val):
        self.src.OPD = val

    @property
    def OPD_no_pupil(self):
        return self.src.OPD_no_pupil

    @OPD_no_pupil.setter
    def OPD_no_pupil(self, val):
        self.src.OPD_no_pupil = val

    def resetOPD(self):
        warning('The use of resetOPD is deprecated. Consider using the Source reset option using the ** '
                'operator to reset the light propagation:\n src**tel')
        self.src ** self

    # ------------------------------------------------------------ operators
    def __mul__(self, obj):
        """Telescope.__mul__ (Telescope.py:524-546)."""
        if type(obj) is list:
            for o in obj:
                self * o
        else:
            self.relay(self.src)
            obj.relay(self.src)
        return self

    def __add__(self, obj):
        if obj.tag == 'atmosphere':
            self.isPaired = True
            self.src * obj * self
            self.atm = obj
        if obj.tag == 'spatialFilter':
            self.spatialFilter = obj
            self * obj
            print('Telescope and Spatial Filter combined!')

    def __sub__(self, obj):
        if obj.tag == 'atmosphere':
            self.isPaired = False
            self.src.reset()
            obj.asterism = None
        if obj.tag == 'spatialFilter':
            self.spatialFilter = None
            print('Telescope and Spatial Filter separated!')

    def print_optical_path(self):
        if self.optical_path is not None:
            print(' ~~> '.join(p[0] for p in self.optical_path))
        else:
            print('No light propagated through the telescope')

    def print_properties(self):
        print(self)

    def properties(self) -> dict:
        self.prop = dict()
        self.prop['diameter'] = f"{'Diameter [m]':<25s}|{self.D:^10.2f}"
        self.prop['resolution'] = f"{'Resolution [px]':<25s}|{self.resolution:^10.0f}"
        self.prop['pixel_size'] = f"{'Pixel size [m]':<25s}|{self.pixelSize:^10.2f}"
        self.prop['surface'] = f"{'Surface [m²]':<25s}|{self.pixelArea * self.pixelSize ** 2:^10.2f}"
        self.prop['obstruction'] = f"{'Central obstruction [%]':<25s}|{self.centralObstruction * 100:^10.0f}"
        self.prop['n_pix_pupil'] = f"{'Pixels in pupil':<25s}|{self.pixelArea:^10.0f}"
        self.prop['fov'] = f"{'Field of view [arcsec]':<25s}|{self.fov:^10.2f}"
        return self.prop

    def __repr__(self):
        self.properties()
        n_char = len(max(self.prop.values(), key=len))
        return (f'\n{" Telescope ":-^{n_char}}\n' + ''.join(v + '\n' for v in self.prop.values())
                + f'{"":-^{n_char}}\n')
