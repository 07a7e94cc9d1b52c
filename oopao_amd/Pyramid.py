"""Pyramid: modulated pyramid WFS drop-in (Pyramid.py:25-1230).

Same constructor, attributes, setters and relay protocol as the reference.
Every frame is computed on the GPU by the HIP library (oopao_amd/csrc/pwfs.hip):
the modulation loop, the FFT chain, the |E|^2 accumulation and the detector
binning run as two fused kernels plus a reduction, and the slopesMaps signal
by one more kernel. Numbers are float32/complex64; public arrays are float64
numpy copies in the reference's layout.

There is no CPU fallback: unsupported geometries raise OopaoError.
"""
import numpy as np
import torch

from . import _lib
from ._device import to_dev, to_host
from .Detector import Detector
from .pwfs_engine import PWFSEngine, SlopesEngine
from .tools import OopaoError, WorkspaceAllocationError, warning


class Pyramid:
    def __init__(self, nSubap: float, telescope, modulation: float, lightRatio: float,
                 postProcessing: str = 'slopesMaps', psfCentering: bool = True, n_pix_separation: float = 2.,
                 calibModulation: float = 50., n_pix_edge: float = None, extraModulationFactor: int = 0,
                 binning: int = 1, nTheta_user_defined: int = None, userValidSignal: bool = None,
                 old_mask: bool = False, rooftop: float = 0, theta_rotation: float = 0, delta_theta: float = 0.,
                 user_modulation_path: list = None, pupilSeparationRatio: float = None, edgePixel: int = None,
                 zeroPadding: int = None, nchunk: int = None, pts_per_task: int = None,
                 max_batch_bytes: float = 8e9, fractional_sampling: bool = False):
        self.gpu_available = False          # batching stays enabled in InteractionMatrix
        self.nJobs = 1
        self.precision = np.float32
        self.precision_complex = np.complex64
        self.telescope = telescope
        self.device = telescope.device
        # fractional_sampling (no counterpart in the reference, which raises here): a non-integer number of
        # pupil pixels per subaperture, the detector integrating the unbinned frame by area
        self.fractional_sampling = bool(fractional_sampling)
        if (self.telescope.resolution / nSubap) % 2 != 0 and self.telescope.resolution / nSubap != 1 \
                and not self.fractional_sampling:
            raise OopaoError('The resolution should be an even number and be a multiple of 2**i where i>=2')
        if self.telescope.src is None:
            raise OopaoError('The telescope was not coupled to any source object! Make sure to couple it with '
                             'an src object using src*tel')
        self.src = self.telescope.src
        self.wavelength_calibration = self.src.wavelength
        self.delta_theta = delta_theta
        self.nTheta_user_defined = nTheta_user_defined
        self.extraModulationFactor = extraModulationFactor
        self.nSubap = nSubap
        self.edgePixel = n_pix_edge
        self.centerPixel = 0
        self.postProcessing = postProcessing
        self.userValidSignal = userValidSignal
        self.psfCentering = psfCentering
        self.binning = binning
        self.old_mask = old_mask
        self.user_modulation_path = user_modulation_path
        self.modulation_length = 2 * np.pi
        self.pupilSeparationRatio = pupilSeparationRatio
        self._weight_vector = None
        if edgePixel is not None:
            raise OopaoError('The use of the edgePixel property has been deprecated. Consider using n_pix_edge instead')
        if pupilSeparationRatio is not None:
            raise OopaoError('The use of the pupilSeparationRatio property has been deprecated. Consider using '
                             'n_pix_separation instead')
        self.n_pix_separation = n_pix_separation
        self.sx = [0, 0, 0, 0]
        self.sy = [0, 0, 0, 0]
        if n_pix_edge is None:
            self.n_pix_edge = self.n_pix_separation // 2
        else:
            self.n_pix_edge = n_pix_edge
            if n_pix_edge != self.n_pix_separation // 2:
                warning('The recommanded value for n_pix_edge is ' + str(self.n_pix_separation // 2)
                        + ' instead of ' + str(n_pix_edge))
        self.joblib_setting = 'threads'
        self.rooftop = rooftop
        self.theta_rotation = theta_rotation
        if zeroPadding is not None:
            raise OopaoError('The use of the zeroPadding property has been deprecated')
        n = self.telescope.resolution
        # geometry (Pyramid.py:276-289)
        full = (self.nSubap * 2 + self.n_pix_separation + self.n_pix_edge * 2) * n / self.nSubap
        if self.fractional_sampling and abs(full - round(full)) > 1e-9 * full:
            raise OopaoError('fractional_sampling needs an integer padded resolution; '
                             '(2 nSubap + n_pix_separation + 2 n_pix_edge) n / nSubap = %g' % full)
        self.resolution = int(round(full)) if self.fractional_sampling else int(full)
        self.zeroPaddingFactor = self.resolution / n
        self.zeroPadding = (self.resolution - n) // 2
        self.tag = 'pyramid'
        self.cam = Detector(round(nSubap * self.zeroPaddingFactor), device=self.device)
        self.focal_plane_camera = Detector(int((modulation * 4 + 12) * self.zeroPaddingFactor * n /
                                               self.telescope.initial_resolution),
                                           psf_sampling=self.zeroPaddingFactor * n / self.telescope.initial_resolution,
                                           device=self.device)
        self.focal_plane_camera.is_focal_plane_camera = True
        self._lightRatio = lightRatio
        self.calibModulation = n / 2 - 1 if calibModulation >= n / 2 else calibModulation
        self.isInitialized = False
        self.isCalibrated = False
        self._delta_Tip = 0
        self._delta_Tilt = 0
        self.center = self.resolution // 2
        self._spatialFilter = None
        self.rad2arcsec = (180 / np.pi) * 3600
        self.fov = self.rad2arcsec * self.resolution / self.zeroPaddingFactor * (self.src.wavelength / self.telescope.D)
        self.fov_l_d = self.resolution / self.zeroPaddingFactor
        self.max_fov_arcsec = self.fov / 2
        self.n_max = 1e9
        self.max_batch_bytes = max_batch_bytes
        # modulation tip/tilt (Pyramid.py:334-339): Tip varies along x (columns), Tilt along y (rows)
        lin = np.linspace(-np.pi, np.pi, n) * (n / self.telescope.initial_resolution)
        self._tip_lin = lin
        self._tilt_lin = lin
        # HIP engine
        self.engine = PWFSEngine(n, self.resolution, self.cam.resolution, self.device, nchunk=nchunk,
                                 pts_per_task=pts_per_task)
        # (opd, wavelength, amplitude) of the last measurement, for the focal-plane camera
        self._last_opd = None
        self._last_weighted = False
        self._raw_cache = None
        self.mask_computation()
        self._setup_slopes()
        self.slopesUnits = 1
        self.referenceSignal = 0
        self._referenceSignal_2D = None
        self.referencePyramidFrame = 0
        self._signal = None
        self._signal_2D = None
        self.modulation = modulation
        print('Selection of the valid pixels...')
        self.initialization()
        print('Acquisition of the reference slopes and units calibration...')
        self.modulation = modulation
        self.wfs_calibration()
        self.src ** self.telescope
        self.wfs_measure()
        print(self)

    # ------------------------------------------------------------------ mask
    def get_phase_mask(self, resolution, n_subap, n_pix_separation, n_pix_edge, psf_centering=False,
                       sx=[0, 0, 0, 0], sy=[0, 0, 0, 0]):
        """Pyramid phase mask (Pyramid.py:453-480)."""
        rooftop_pixels = self.rooftop / np.sqrt(2)
        n_tot = int((n_subap * 2 + n_pix_separation + n_pix_edge * 2) * self.telescope.resolution / self.nSubap)
        norma = self.telescope.resolution / self.nSubap
        lim = np.pi
        x = np.linspace(-(1 - self.psfCentering * 1 / n_tot) * lim, (1 - self.psfCentering * 1 / n_tot) * lim, n_tot,
                        endpoint=self.psfCentering)
        x_, y_ = np.meshgrid(x, x)
        xr = x_ * np.cos(self.theta_rotation) - y_ * np.sin(self.theta_rotation)
        yr = y_ * np.cos(self.theta_rotation) + x_ * np.sin(self.theta_rotation)
        r = (self.nSubap + self.n_pix_separation) / 2
        P1 = xr * r + x_ * sx[0] + yr * r - y_ * sy[0] + rooftop_pixels
        P2 = -xr * r + x_ * sx[1] + yr * r - y_ * sy[1]
        P3 = -xr * r + x_ * sx[2] - yr * r - y_ * sy[2] + rooftop_pixels
        P4 = xr * r + x_ * sx[3] - yr * r - y_ * sy[3]
        return -np.max(np.stack([P1, P2, P3, P4]) * norma, axis=0)

    def get_phase_mask_old(self, resolution, n_subap, n_pix_separation, n_pix_edge, psf_centering=False,
                           sx=[0, 0, 0, 0], sy=[0, 0, 0, 0]):
        """Previous pyramid mask: one zero-mean tip/tilt ramp per quadrant (Pyramid.py:482-539)."""
        n_tot = int((n_subap * 2 + n_pix_separation + n_pix_edge * 2) * self.telescope.resolution / self.nSubap)
        norma = (self.telescope.resolution / self.nSubap) / 4
        h = n_tot // 2
        L = n_subap + n_pix_separation
        m = np.zeros([n_tot, n_tot])

        def ramps(nx, ny, limx, limy, endx, endy):
            tip, tilt = np.meshgrid(np.linspace(-limx, limx, nx, endpoint=endx),
                                    np.linspace(-limy, limy, ny, endpoint=endy))
            return tip - tip.mean(), tilt - tilt.mean()
        if psf_centering:
            tip, tilt = ramps(h, h, np.pi, np.pi, False, False)
            m[:h, :h] = (tip * (L + sx[0]) + tilt * (L - sy[0])) * norma
            m[:h, -h:] = (-tip * (L - sx[1]) + tilt * (L - sy[1])) * norma
            m[-h:, -h:] = (-tip * (L - sx[2]) - tilt * (L + sy[2])) * norma
            m[-h:, :h] = (tip * (L + sx[3]) - tilt * (L + sy[3])) * norma
        else:
            lp = np.pi
            lm = np.pi - 2 * np.pi / n_tot
            tip1, tilt1 = ramps(h + 1, h + 1, lp, lp, True, True)
            tip2, tilt2 = ramps(h + 1, h - 1, lp, lm, True, False)
            tip3, tilt3 = ramps(h - 1, h - 1, lm, lm, False, False)
            tip4, tilt4 = ramps(h - 1, h + 1, lm, lp, False, True)
            m[:h + 1, :h + 1] = (tip1 * (L + sx[0]) + tilt1 * (L - sy[0])) * norma
            m[:h + 1, -h + 1:] = (-tip4 * (L - sx[1]) + tilt4 * (L - sy[1])) * norma
            m[-h + 1:, -h + 1:] = (-tip3 * (L - sx[2]) - tilt3 * (L + sy[2])) * norma
            m[-h + 1:, :h + 1] = (tip2 * (L + sx[3]) - tilt2 * (L + sy[3])) * norma
        return -m

    def _phase_mask(self, sx, sy):
        f = self.get_phase_mask_old if self.old_mask else self.get_phase_mask
        return f(self.resolution, self.nSubap, self.n_pix_separation, self.n_pix_edge, self.psfCentering, sx, sy)

    def _apply_mask(self):
        """Send the current mask to the HIP engine. With a spatial filter the mask holds it
        (mask = initial_mask * filter, Pyramid.py:1072) and, without PSF centring, the focal field is
        multiplied by it once more before the mask (Pyramid.py:633-636)."""
        sf = self._spatialFilter
        if sf is None:
            self.engine.set_mask(self.m, self.mask)
        else:
            factor = None if self.psfCentering else np.asarray(sf)
            self.engine.set_mask(self.m, self.mask, factor=factor, general=True)

    def mask_computation(self):
        """Pyramid.mask_computation (Pyramid.py:368-391)."""
        print('Pyramid Mask initialization...')
        self.m = self._phase_mask(self.sx, self.sy)
        self.initial_m = self.m.copy()
        self.mask = np.complex64(np.exp(1j * self.m))
        self.initial_mask = np.copy(self.mask)
        self._apply_mask()

    def apply_shift_wfs(self, sx=None, sy=None, mis_reg=None, units='pixels'):
        """Per-quadrant pupil shifts (Pyramid.py:393-451). Uniform shifts keep the mask separable;
        others run the general HIP path."""
        sx = 0 if sx is None else sx
        sy = 0 if sy is None else sy
        if mis_reg is not None:
            sx = [mis_reg.dX_1, mis_reg.dX_2, mis_reg.dX_3, mis_reg.dX_4]
            sy = [mis_reg.dY_1, mis_reg.dY_2, mis_reg.dY_3, mis_reg.dY_4]
        f = 2 if self.old_mask else 1     # the previous mask's shift convention (Pyramid.py:401-405)
        factor = f if units == 'pixels' else f / (self.telescope.D / self.nSubap)
        if np.isscalar(sx) and np.isscalar(sy):
            shift_x = [factor * sx] * 4
            shift_y = [factor * sy] * 4
        else:
            if len(sx) == 4 and len(sy) == 4:
                shift_x = [v * factor for v in sx]
                shift_y = [v * factor for v in sy]
            else:
                raise OopaoError('Wrong size for sx and/or sy, a list of 4 values is expected.')
        if np.max(np.abs(shift_x)) > self.n_pix_edge or np.max(np.abs(shift_y)) > self.n_pix_edge:
            warning('The Pyramid pupils have been shifted outside of the detector!' +
                    'Wrapping of the signal is currently occuring!!')
        self.sx = np.asarray(shift_x) / factor
        self.sy = np.asarray(shift_y) / factor
        self.m = self._phase_mask(shift_x, shift_y)
        self.mask = np.complex64(np.exp(1j * self.m))
        self._apply_mask()
        self.slopesUnits = 1
        self.referenceSignal = 0
        self._referenceSignal_2D = None
        self.wfs_calibration()

    # ---------------------------------------------------------- slopes setup
    def _quadrant_origins(self):
        """Row/column origins of quadrants 1..4 (Pyramid.py:946-961)."""
        n = self.telescope.resolution
        e = int(np.round((self.n_pix_separation / self.nSubap) * n / (n / self.nSubap) / 2 / self.binning))
        c = int(np.round((self.cam.resolution / self.binning) / 2))
        w = int(np.ceil(self.nSubap / self.binning))
        rows = [c - e - w, c - e - w, c + e, c + e]
        cols = [c - e - w, c + e, c + e, c - e - w]
        return rows + cols, w

    def _setup_slopes(self):
        quad, w = self._quadrant_origins()
        self._w = w
        self.slopes_engine = SlopesEngine(self.device, self.cam.resolution // self.binning, w, quad)

    # ------------------------------------------------------------ modulation
    @property
    def modulation(self):
        return self._modulation

    @modulation.setter
    def modulation(self, val):
        """Modulation setter (Pyramid.py:1107-1154)."""
        self._modulation = val
        if self._modulation >= (self.telescope.resolution // 2):
            raise OopaoError('Error the modulation radius is too large for this resolution!' +
                             'Consider using a larger telescope resolution!')
        if val != 0 or self.user_modulation_path is not None:
            if self.user_modulation_path is not None:
                self.modulation_path = list(self.user_modulation_path)
                self.nTheta = len(self.user_modulation_path)
            else:
                perimeter = self.modulation_length * self._modulation
                if self.nTheta_user_defined is None:
                    self.nTheta = 4 * int((self.extraModulationFactor + np.ceil(perimeter / 4)))
                else:
                    self.nTheta = self.nTheta_user_defined
                self.thetaModulation = np.linspace(0 + self.delta_theta, 2 * np.pi + self.delta_theta,
                                                   self.nTheta, endpoint=False)
                self.modulation_path = [[self.modulation * np.cos(t) + self.delta_Tip,
                                         self.modulation * np.sin(t) + self.delta_Tilt] for t in self.thetaModulation]
            path = np.asarray(self.modulation_path, dtype=np.float64)
        else:
            self.nTheta = 1
            path = np.zeros((1, 2))
        self._path = path
        self.engine.set_modulation(path, self._tip_lin, self._tilt_lin, self.psfCentering)
        if self.isCalibrated:
            print('Updating the reference slopes')
            self.slopesUnits = 1
            self.referenceSignal = 0
            self._referenceSignal_2D = None
            self.wfs_calibration()
            print('Done!')

    @property
    def phaseBuffModulationLowres(self):
        """float32 tip/tilt phase per modulation point [T, n, n] (Pyramid.py:1133-1139)."""
        tip, tilt = np.meshgrid(self._tip_lin, self._tilt_lin)
        return np.stack([(p[0] * tip + p[1] * tilt).astype(np.float32) for p in self._path])

    @property
    def delta_Tip(self):
        return self._delta_Tip

    @delta_Tip.setter
    def delta_Tip(self, val):
        self._delta_Tip = val
        if self.isCalibrated:
            self.modulation = self.modulation

    @property
    def delta_Tilt(self):
        return self._delta_Tilt

    @delta_Tilt.setter
    def delta_Tilt(self, val):
        self._delta_Tilt = val
        if self.isCalibrated:
            self.modulation = self.modulation

    # ---------------------------------------------------------- valid pixels
    def grabQuadrant(self, n, cameraFrame=None):
        quad, w = self._quadrant_origins()
        frame = self.cam.frame if cameraFrame is None else np.asarray(cameraFrame)
        r, c = quad[n - 1], quad[4 + n - 1]
        return frame[r:r + w, c:c + w].copy()

    def grabFullQuadrant(self, n, cameraFrame=None):
        n_tot = self.cam.resolution
        frame = self.cam.frame if cameraFrame is None else np.asarray(cameraFrame)
        if n == 3:
            return frame[:n_tot // 2, :n_tot // 2]
        if n == 4:
            return frame[:n_tot // 2, -n_tot // 2:]
        if n == 1:
            return frame[-n_tot // 2:, -n_tot // 2:]
        if n == 2:
            return frame[-n_tot // 2:, :n_tot // 2]

    def initialization(self):
        """Valid pixel selection at the calibration modulation (Pyramid.py:541-589)."""
        self.src ** self.telescope
        if self.userValidSignal is None:
            if self.lightRatio == 0:
                self.cam.frame = np.ones([self.cam.resolution, self.cam.resolution])
            else:
                print('The valid pixel are selected on flux considerations')
                self.modulation = self.calibModulation
                self.wfs_measure()
            self.initFrame = self.cam.frame
            if self.postProcessing[:10] == 'slopesMaps':
                I4Q = sum(self.grabQuadrant(k) for k in (1, 2, 3, 4))
                self.I4Q = I4Q
                self.validI4Q = (self.I4Q >= self.lightRatio * self.I4Q.max())
                self.validSignal = np.concatenate((self.validI4Q, self.validI4Q))
                self.nSignal = int(np.sum(self.validSignal))
            if self.postProcessing[:9] == 'fullFrame':
                self.validSignal = (self.initFrame >= self.lightRatio * self.initFrame.max())
                self.nSignal = int(np.sum(self.validSignal))
        else:
            print('You are using a user-defined mask for the selection of the valid pixel')
            if self.postProcessing[:10] == 'slopesMaps':
                self.validI4Q = np.asarray(self.userValidSignal, bool)
                self.validSignal = np.concatenate((self.validI4Q, self.validI4Q))
            if self.postProcessing[:9] == 'fullFrame':
                self.validSignal = np.asarray(self.userValidSignal, bool)
            self.nSignal = int(np.sum(self.validSignal))
        if self.postProcessing[:10] == 'slopesMaps':
            self.slopes_engine.set_valid(self.validI4Q)
        else:
            self._valid_full = torch.from_numpy(self.validSignal.astype(bool)).to(self.device)
        self.isInitialized = True

    @property
    def lightRatio(self):
        return self._lightRatio

    @lightRatio.setter
    def lightRatio(self, val):
        self._lightRatio = val
        if getattr(self, 'isInitialized', False):
            print('Updating the map if valid pixels ...')
            self.validI4Q = (self.I4Q >= self._lightRatio * self.I4Q.max())
            self.validSignal = np.concatenate((self.validI4Q, self.validI4Q))
            self.validPix = (self.initFrame >= self.lightRatio * self.initFrame.max())
            if self.postProcessing[:10] == 'slopesMaps':
                self.nSignal = int(np.sum(self.validSignal))
                self.slopes_engine.set_valid(self.validI4Q)
                if self._referenceSignal_2D is not None:
                    self.slopes_engine.set_reference(self._referenceSignal_2D)
            if self.postProcessing[:9] == 'fullFrame':
                self.nSignal = int(np.sum(self.validPix))
            print('Done!')

    # ------------------------------------------------------------ calibration
    def wfs_calibration(self, src=None):
        """Reference signal on the flat wavefront (Pyramid.py:591-610)."""
        tmp = (self.src._OPD.clone() if self.src._OPD is not None else None,
               self.src._OPD_no_pupil.clone() if self.src._OPD_no_pupil is not None else None)
        if src is None:
            self.src ** self.telescope
        else:
            self.src = src
        self.relay(self.src)
        self._ref_frame = (self._last_opd, self._last_weighted)
        self._sync_reference()
        s2d, s = self._signal_processing_dev()
        self._referenceSignal_2D = s2d[0] if s2d.dim() == 3 else s2d
        self.referenceSignal = to_host(s[0] if s.dim() == 2 else s)
        if self.postProcessing[:10] == 'slopesMaps':
            self.slopes_engine.set_reference(self._referenceSignal_2D)
        if self.isCalibrated is False:
            print('WFS calibrated!')
        self.isCalibrated = True
        self.src._OPD, self.src._OPD_no_pupil = tmp

    def _sync_reference(self):
        """Make the slopes kernel subtract the current referenceSignal_2D (zeros when reset)."""
        if self.postProcessing[:10] != 'slopesMaps' or not hasattr(self.slopes_engine, 'valid'):
            return
        if self._referenceSignal_2D is None:
            self.slopes_engine.set_reference(torch.zeros((2 * self._w, self._w), device=self.device))
        else:
            self.slopes_engine.set_reference(self._referenceSignal_2D)

    @property
    def slopesUnits(self):
        return self._slopesUnits

    @slopesUnits.setter
    def slopesUnits(self, val):
        self._slopesUnits = val
        if hasattr(self, 'slopes_engine'):
            self.slopes_engine.units = float(val)

    @property
    def referenceSignal_2D(self):
        return 0 if self._referenceSignal_2D is None else to_host(self._referenceSignal_2D)

    @referenceSignal_2D.setter
    def referenceSignal_2D(self, val):
        if np.isscalar(val):
            self._referenceSignal_2D = None
            self.slopes_engine.set_reference(torch.zeros((2 * self._w, self._w), device=self.device))
        else:
            self._referenceSignal_2D = torch.as_tensor(np.asarray(val), dtype=torch.float32, device=self.device)
            self.slopes_engine.set_reference(self._referenceSignal_2D)

    # ------------------------------------------------------------ measurement
    def _amp(self):
        """Field amplitude sqrt(src.intensity / nTheta) on the device (Pyramid.py:674)."""
        return torch.sqrt(self.src._intensity / self.nTheta).contiguous()

    def _batch_size(self, B):
        per = max(1, self.engine.bytes_per_screen())
        return max(1, min(B, int(self.max_batch_bytes // per)))

    def slope_bands(self):
        """Detector-column intervals the slopesMaps signal reads (the quadrants of grabQuadrant,
        Pyramid.py:838-853, 946-961), for a slopes-only frame; None when a measurement needs the whole
        frame (another post-processing mode, the general path, detector binning)."""
        if (self.postProcessing != 'slopesMaps' or getattr(self, 'engine', None) is None
                or self.engine.mode != 'fast' or self.engine.fractional or self.binning != 1):
            return None
        quad, w = self._quadrant_origins()
        return [(c, c + w) for c in sorted(set(quad[4:]))]

    def frames_dev(self, opd, weighted=False, bands=None, amp=None, scale=None):
        """Detector frames for device OPD maps [B, n, n] (metres) of the current source. ``weighted``
        applies weight_vector to the modulation points (the 2-D measurement, Pyramid.py:750-755);
        ``bands`` restricts the frames to those detector columns (see slope_bands)."""
        amp = self._amp() if amp is None else amp
        scale = 2 * np.pi / self.src.wavelength if scale is None else scale
        B = opd.shape[0]
        bs = self._batch_size(B)
        while True:
            try:
                if bs >= B:
                    return self.engine.frames(opd.contiguous(), scale, amp, weighted=weighted, bands=bands)
                out = torch.empty((B, self.engine.cam, self.engine.cam), dtype=torch.float32, device=self.device)
                for j in range(0, B, bs):
                    self.engine.frames(opd[j:j + bs].contiguous(), scale, amp, out=out[j:j + bs],
                                       weighted=weighted, bands=bands)
                return out
            except WorkspaceAllocationError as e:
                # the reference frees memory by measuring fewer screens at a time (Pyramid.py:713-740):
                # halve the batch, keep the smaller size for later calls, and measure again
                if bs == 1:
                    raise
                bs = max(1, min(bs, B) // 2)
                self.max_batch_bytes = bs * max(1, self.engine.bytes_per_screen())
                warning('%s; measuring %d screens at a time' % (e.string, bs))

    def wfs_measure(self, phase_in=None, integrate=True):
        """Pyramid.wfs_measure (Pyramid.py:666-833): all modulation points, 2-D or batched 3-D."""
        if self.isInitialized and self.isCalibrated:
            if self.wavelength_calibration != self.src.wavelength:
                raise OopaoError('A change in wavelength was detected in the WFS object \n' +
                                 'Make sure that the correct source is propagated in the WFS object or '
                                 're-calibrate with the correct source.')
        if phase_in is not None:
            self.src.phase = phase_in
        self._stacked_sources = False
        opd = self.src._OPD
        batched = opd.dim() == 3
        weighted = (not batched) and self._weight_vector is not None
        opd_b = opd if batched else opd[None]
        amp = self._amp()
        scale = 2 * np.pi / self.src.wavelength
        # slope-only frame: when the signal is all this measurement produces and the camera readout is
        # an identity, compute only the quadrant columns; cam.frame computes the whole frame when read
        bands = None
        if integrate and self.isInitialized and self.isCalibrated and self.cam.identity_readout():
            bands = self.slope_bands()
        frames = self.frames_dev(opd_b, weighted=weighted, bands=bands, amp=amp, scale=scale)
        self._last_opd = (opd_b, self.src.wavelength, amp)
        self._last_weighted = weighted
        self._raw_cache = None
        if self.binning != 1:
            b = self.binning
            c = frames.shape[-1]
            frames = frames.reshape(-1, c // b, b, c // b, b).sum(dim=(2, 4))
        self.cam._integrated_time += self.telescope.samplingTime
        self.cam.integrate(frames if batched else frames[0])
        if bands is not None:
            opd_keep = opd_b.clone()

            def whole_frame():
                f = self.frames_dev(opd_keep, weighted=weighted, amp=amp, scale=scale)
                return f if batched else f[0]
            self.cam._defer_full(whole_frame)
        if integrate and self.isInitialized and self.isCalibrated:
            s2d, s = self._signal_processing_dev()
            self._signal_2D, self._signal = s2d, s
            self._batched = batched

    def wfs_integrate(self):
        if self.isInitialized and self.isCalibrated:
            s2d, s = self._signal_processing_dev()
            return to_host(s2d), to_host(s)
        return None, None

    def _signal_processing_dev(self):
        frame = self.cam._frame
        if self.postProcessing == 'slopesMaps':
            s2d, s = self.slopes_engine.slopes(frame)
            return (s2d, s) if frame.dim() == 3 else (s2d[0], s[0])
        return self._signal_processing_other(frame)

    def _signal_processing_other(self, frame):
        """Other post-processing modes (Pyramid.py:856-926) as device tensor expressions."""
        f = frame if frame.dim() == 3 else frame[None]
        pp = self.postProcessing
        ref = self._referenceSignal_2D if self._referenceSignal_2D is not None else 0
        if pp.startswith('slopesMaps'):
            quad, w = self._quadrant_origins()
            v = torch.from_numpy(self.validI4Q.astype(np.float32)).to(self.device)
            Iq = [f[:, quad[k]:quad[k] + w, quad[4 + k]:quad[4 + k] + w] * v for k in range(4)]
            if pp == 'slopesMaps_incidence_flux':
                norma = float(self.src.nPhoton * self.telescope.samplingTime * (self.telescope.D / self.nSubap) ** 2)
                norma = torch.full((f.shape[0], 1, 1), norma, device=self.device)
            elif pp == 'slopesMaps_camera_flux':
                norma = f.double().mean(dim=(1, 2)).float()[:, None, None]
            else:
                raise OopaoError('unknown postProcessing ' + pp)
            Sx = Iq[0] - Iq[1] + Iq[3] - Iq[2]
            Sy = Iq[0] - Iq[3] + Iq[1] - Iq[2]
            s2d = (torch.cat((Sx, Sy), dim=1) / norma - ref) * self.slopesUnits
            vs = torch.from_numpy(self.validSignal.astype(bool)).to(self.device)
            s = s2d[:, vs]
        else:
            if pp == 'fullFrame_camera_flux':
                norma = f.double().mean(dim=(1, 2)).float()[:, None, None]
            elif pp == 'fullFrame_incidence_flux':
                norma = float(self.src.nPhoton * self.telescope.samplingTime * (self.telescope.D / self.nSubap) ** 2) / 4
                norma = torch.full((f.shape[0], 1, 1), norma, device=self.device)
            elif pp == 'fullFrame_sum_flux':
                norma = f.double().sum(dim=(1, 2)).float()[:, None, None]
            elif pp == 'fullFrame':
                norma = f[:, self._valid_full].double().sum(dim=1).float()[:, None, None]
            else:
                raise OopaoError('unknown postProcessing ' + pp)
            s2d = f / norma - ref
            s = s2d[:, self._valid_full]
        if frame.dim() == 2:
            return s2d[0], s[0]
        return s2d, s

    @property
    def norma(self):
        """Normalisation of the last signal (the value signalProcessing stores, Pyramid.py:838-926),
        computed from the current camera frame; one value per frame for a batched measurement."""
        frame = self.cam._frame
        if frame is None:
            return None
        f = (frame if frame.dim() == 3 else frame[None]).double()
        pp = self.postProcessing
        sub_flux = float(self.src.nPhoton * self.telescope.samplingTime * (self.telescope.D / self.nSubap) ** 2)
        if pp == 'slopesMaps':
            quad, w = self._quadrant_origins()
            i4q = sum(f[:, quad[k]:quad[k] + w, quad[4 + k]:quad[4 + k] + w] for k in range(4))
            v = torch.from_numpy(np.asarray(self.validI4Q, dtype=bool)).to(self.device)
            out = i4q[:, v].mean(dim=1)
        elif pp in ('slopesMaps_camera_flux', 'fullFrame_camera_flux'):
            out = f.mean(dim=(1, 2))
        elif pp == 'slopesMaps_incidence_flux':
            out = torch.full((f.shape[0],), sub_flux, dtype=torch.float64)
        elif pp == 'fullFrame_incidence_flux':
            out = torch.full((f.shape[0],), sub_flux / 4, dtype=torch.float64)
        elif pp == 'fullFrame_sum_flux':
            out = f.sum(dim=(1, 2))
        elif pp == 'fullFrame':
            out = f[:, self._valid_full].sum(dim=1)
        else:
            raise OopaoError('unknown postProcessing ' + pp)
        out = out.cpu().numpy()
        return float(out[0]) if frame.dim() == 2 else out

    def signalProcessing(self, cameraFrame=None):
        """(signal_2D, signal) of the current camera frame (Pyramid.py:835-926)."""
        s2d, s = self._signal_processing_dev()
        return to_host(s2d), self._sig_host(s)

    def _sig_host(self, s):
        if getattr(self, '_stacked_sources', False):       # asterism: [nSrc, nSig] (Pyramid.py:1171-1173)
            return s.detach().double().cpu().numpy()
        a = to_host(s)
        if a is not None and a.ndim == 2:  # [B, nSig] -> [nSig, B]
            a = a.T.copy()
        return a

    # --------------------------------------------------------------- outputs
    @property
    def signal(self):
        return self._sig_host(self._signal)

    @signal.setter
    def signal(self, val):
        self._signal = None if val is None else torch.as_tensor(np.asarray(val), dtype=torch.float32,
                                                                 device=self.device)

    @property
    def pyramidSignal(self):
        return self.signal

    @property
    def signal_2D(self):
        if getattr(self, '_stacked_sources', False):       # asterism: [nSrc, 2w, w]
            return self._signal_2D.detach().double().cpu().numpy()
        return to_host(self._signal_2D)

    @property
    def pyramidSignal_2D(self):
        return self.signal_2D

    @property
    def validSignal(self):
        return self._validSignal

    @validSignal.setter
    def validSignal(self, val):
        self._validSignal = val
        self.valid_signal_2D = val

    @property
    def resolution(self):
        return self._resolution

    @resolution.setter
    def resolution(self, val):
        self._resolution = val
        self.nRes = val

    @property
    def spatialFilter(self):
        return self._spatialFilter

    @spatialFilter.setter
    def spatialFilter(self, val):
        """Spatial filter setter (Pyramid.py:1055-1085): the mask becomes initial_mask * filter."""
        self._spatialFilter = val
        if not self.isInitialized:
            return
        if val is None:
            print('No spatial filter considered')
            self.mask = self.initial_mask
        elif np.shape(val) == self.mask.shape:
            print('A spatial filter is now considered')
            self.mask = self.initial_mask * np.asarray(val)
        else:
            warning('Wrong shape for the spatial filter. No spatial filter attached to the mask')
            self.mask = self.initial_mask
        self._apply_mask()
        if self.isCalibrated:
            print('Updating the reference slopes and Wavelength Calibration for the new modulation...')
            self.slopesUnits = 1
            self.referenceSignal = 0
            self._referenceSignal_2D = None
            self.wfs_calibration()
            print('Done!')

    def _spatial_filter_fields(self):
        """Focal field of the unmodulated source and its filtered pupil-plane image (the diagnostics
        em_field_spatial_filter / pupil_plane_spatial_filter of Pyramid.py:676-684), on the device."""
        if self._spatialFilter is None or self._last_opd is None:
            return None, None
        opd, wavelength, amp = self._last_opd
        n, N = self.engine.n, self.engine.N
        c0 = N // 2 - n // 2
        field = amp * torch.exp(1j * opd[0] * (2 * np.pi / wavelength))
        support = torch.zeros((N, N), dtype=torch.complex64, device=self.device)
        support[c0:c0 + n, c0:c0 + n] = field
        xx = torch.arange(N, device=self.device, dtype=torch.float64)
        ph = torch.exp(-1j * np.pi * (N + 1) / N * (xx[None, :] + xx[:, None])).to(torch.complex64)
        em = torch.fft.fft2(support * ph)
        sf = torch.as_tensor(np.asarray(self._spatialFilter), device=self.device).to(torch.complex64)
        return em, torch.fft.ifft2(em * sf)

    @property
    def em_field_spatial_filter(self):
        return to_host(self._spatial_filter_fields()[0])

    @property
    def pupil_plane_spatial_filter(self):
        return to_host(self._spatial_filter_fields()[1])

    @property
    def weight_vector(self):
        return self._weight_vector

    @weight_vector.setter
    def weight_vector(self, val):
        """Per-point weights of the 2-D measurement: raw = sum_i w_i I_i / nTheta (Pyramid.py:750-755)."""
        self._weight_vector = None if val is None else np.asarray(val, np.float64).reshape(-1)
        if hasattr(self, 'engine'):
            self.engine.set_weights(self._weight_vector)

    # ---------------------------------------------------------------- relay
    def relay(self, src):
        """Pyramid.relay (Pyramid.py:1156-1176)."""
        if src.tag == 'source':
            src_list = [src]
        elif src.tag == 'asterism':
            src_list = src.src
        sig2d, sig, frames = [], [], []
        for s in src_list:
            s.optical_path.append([self.tag, self])
            self.src = s
            self.wfs_measure()
            sig2d.append(self._signal_2D)
            sig.append(self._signal)
            frames.append(self.cam._frame_getter())
        if len(src_list) > 1:
            self._signal_2D = torch.stack(sig2d) if sig2d[0] is not None else None
            self._signal = torch.stack(sig) if sig[0] is not None else None
        self._stacked_sources = len(src_list) > 1
        self._frame_getters = frames

    @property
    def frames(self):
        """Camera frames of the sources of the last relay (Pyramid.py:1170-1175), host float64."""
        getters = getattr(self, '_frame_getters', None)
        if not getters:
            return None
        return np.squeeze(np.array([to_host(g()) for g in getters]))

    def focal_intensity_dev(self, window):
        """Sum over modulation points (and over the screens of a batched measurement) of the focal-plane
        intensity |fft2(E_i)|^2 / N^2, on the centred window x window crop of the last measured field
        (Pyramid.py:642 modulation_camera_em, 1183-1190)."""
        if self._last_opd is None:
            raise OopaoError('no field has been propagated to the WFS yet')
        opd, wavelength, amp = self._last_opd
        I = self.engine.focal(opd, 2 * np.pi / wavelength, amp, window)
        return I.sum(dim=0)

    def __mul__(self, obj):
        """Pyramid.__mul__ (Pyramid.py:1178-1203): WFS camera or modulation focal-plane camera."""
        if obj.tag == 'detector':
            if getattr(obj, 'is_focal_plane_camera', False):
                obj._integrated_time += self.telescope.samplingTime
                if obj.resolution > self.resolution:
                    warning('Maximum resolution for focal plane camera is %i, cropping field to this dimension'
                            % self.resolution)
                    window = self.resolution
                else:
                    window = 2 * (obj.resolution // 2)
                frame = self.focal_intensity_dev(window)
                if self.binning != 1:
                    b = self.binning
                    r = obj.resolution // b
                    if frame.shape[0] % r == 0 and frame.shape[1] % r == 0:
                        fy, fx = frame.shape[0] // r, frame.shape[1] // r
                        frame = frame.reshape(r, fy, r, fx).sum(dim=(1, 3))
                    else:
                        warning('The shape of the detector (' + str(tuple(frame.shape)) + ')is not valid with '
                                'the binning value requested:' + str(self.binning) + '! -- Ignoring the binning.')
                obj.integrate(frame)
                return
            obj.integrate(self.cam._resolved())
        else:
            raise OopaoError('Error light propagated to the wrong type of object')

    def get_modulation_frame(self, radius=6, norma=True):
        """Pyramid.get_modulation_frame (Pyramid.py:928-944): the focal-plane camera's frame cropped to
        +-radius lambda/D about its centre, optionally normalised to its maximum."""
        if radius <= 0:
            warning('radius for the field of view must be a strictly positive number. Ignoring the input value.')
            radius = self.telescope.resolution // 2
        self.modulation_camera_frame = np.asarray(self.focal_plane_camera.frame, dtype=float)
        n_trunc = int(self.resolution / 2 - radius * self.zeroPaddingFactor)
        if n_trunc <= 0:
            warning('radius Value is too high as the field of view is limited to ' +
                    str(int(self.fov_l_d / 2)) + ' lambda/D -- ignoring')
            zoom = self.modulation_camera_frame.copy()
        else:
            zoom = self.modulation_camera_frame[n_trunc:-n_trunc, n_trunc:-n_trunc]
        if norma:
            zoom /= zoom.max()
        return zoom

    def _raw_dev(self, args, weighted):
        opd, wavelength, amp = args
        return self.engine.raw(opd[-1:].contiguous(), 2 * np.pi / wavelength, amp, weighted=weighted)[0]

    @property
    def raw_data(self):
        """Unbinned modulated frame [N, N] of the last measurement (of its last screen when batched),
        Pyramid.py:751-757, 820-821; computed on demand (the separable path's unbinned frame, or the
        general path)."""
        if self._last_opd is None:
            return None
        if self._raw_cache is None:
            self._raw_cache = self._raw_dev(self._last_opd, self._last_weighted)
        return to_host(self._raw_cache)

    @property
    def pyramidFrame(self):
        return self.raw_data

    @property
    def referencePyramidFrame(self):
        """Unbinned frame of the calibration wavefront (Pyramid.py:604), computed on first access."""
        v = self._ref_frame
        if isinstance(v, tuple):
            self._ref_frame = to_host(self._raw_dev(*v))
            v = self._ref_frame
        return v

    @referencePyramidFrame.setter
    def referencePyramidFrame(self, val):
        self._ref_frame = val

    def print_properties(self):
        print(self)

    def properties(self) -> dict:
        self.prop = dict()
        self.prop['pupil_diameter'] = f"{'Pupil diameter [px]':<25s}|{self.nSubap:^9d}"
        self.prop['pupil_separation'] = f"{'Pupil separation [px]':<25s}|{self.n_pix_separation:^9.2f}"
        self.prop['fov'] = f"{'Field of view [arcsec]':<25s}|{self.fov:^9.2f}"
        self.prop['modulation'] = f"{'Modulation radius [l/D]':<25s}|{self.modulation:^9.1f}"
        self.prop['psf_sampling'] = f"{'PSF sampling [px/(l/D)]':<25s}|{self.zeroPaddingFactor:^9.2f}"
        self.prop['psf_centering'] = f"{'PSF centering':<25s}|{str(self.psfCentering):^9s}"
        self.prop['n_valid_pixels'] = f"{'Valid pixels':<25s}|{self.nSignal:^9.0f}"
        self.prop['post_processing'] = f"{'Post processing':<25s}|{self.postProcessing:^9s}"
        return self.prop

    def __repr__(self):
        self.properties()
        n_char = len(max(self.prop.values(), key=len))
        return (f'\n{" Pyramid WFS ":-^{n_char}}\n' + ''.join(v + '\n' for v in self.prop.values())
                + f'{"":-^{n_char}}\n')
