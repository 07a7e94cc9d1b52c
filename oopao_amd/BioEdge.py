"""Bi-O Edge wavefront sensor (BioEdge.py:21-1084) on the general HIP PWFS kernels.

The Bi-O Edge replaces the pyramid by four knife-edge amplitude masks (mask_computation,
BioEdge.py:331-390): A = sqrt(BW(x)) with BW a half plane and an optional grey gradient of half
width grey_width (lambda/D), B = sqrt(1 - A^2), C = A^T, D = B^T, each times a tip/tilt phase
for the per-mask pupil shifts. Per modulation point one focal field fft2(support phasor) is
filtered by the four masks and each inverse transform's intensity fills one N x N tile of a
2N x 2N frame (bio_edge_transform, 510-544), the field amplitude being sqrt(0.5) sqrt(I/nTheta).

Here each mask is one general-path PWFS engine (any N; the same row/column transform kernels as
the Pyramid, the mask applied as a full N x N array); the four binned cam x cam images are tiled
into the 2 cam x 2 cam detector frame, which equals binning the tiled 2N x 2N frame by
N / cam = n / nSubap (BioEdge.__mul__, 1033-1055). The quadrants of the slopesMaps processing
are the reference's (grabQuadrant, 840-853), the post-processing modes the Pyramid's.
"""
import numpy as np
import torch

from ._device import to_host
from .Detector import Detector
from .Pyramid import Pyramid
from .pwfs_engine import PWFSEngine, SlopesEngine, Workspace
from .tools import OopaoError, WorkspaceAllocationError, warning


def bioedge_mask_set(N, n, n_initial, nSubap, zero_padding_factor, grey_width=0.0, grey_length=False,
                     sx=(0, 0, 0, 0), sy=(0, 0, 0, 0), quadrants_numbering=(0, 1, 2, 3)):
    """The four complex64 masks [4, N, N] and their tip/tilt phases (BioEdge.py:331-390)."""
    edge = np.zeros(N)
    edge[:N // 2] = 1.0
    r = int(np.round(zero_padding_factor * grey_width * n / n_initial))
    gradient = None
    if grey_width != 0:
        gradient = np.hstack([np.linspace(1, 0.5, r, endpoint=True), np.linspace(0.5, 0, r, endpoint=True)])
        edge[N // 2 - r:N // 2 + r] = gradient
    A = np.sqrt(np.broadcast_to(edge, (N, N)).copy())
    if grey_length is not False:
        rl = int(np.round(zero_padding_factor * grey_length))
        A[:N // 2 - rl, :N // 2] = 1.0
        A[N // 2 + rl:, :N // 2] = 1.0
        A[:N // 2 - rl, N // 2:] = 0.0
        A[N // 2 + rl:, N // 2:] = 0.0
    B = np.sqrt(1 - A ** 2)
    amps = [A, B, A.T.copy(), B.T.copy()]
    ramp = np.linspace(-np.pi, np.pi, N, endpoint=False)
    tip = np.broadcast_to(ramp, (N, N))
    tilt = tip.T
    norma = (nSubap * 2) * (n / nSubap) * (1 / (nSubap * 2))
    sx = [sx] * 4 if np.isscalar(sx) else list(sx)
    sy = [sy] * 4 if np.isscalar(sy) else list(sy)
    if len(sx) == 4 and len(sy) == 4:
        # mask 0 takes the shifts of entry 3, mask 3 those of entry 0 (BioEdge.py:367-371)
        tts = [-sx[3 - k] * tip * norma + sy[3 - k] * tilt * norma for k in range(4)]
    else:
        tts = [np.zeros((N, N))] * 4
    masks = [amps[quadrants_numbering[k]] * np.exp(1j * tts[k]) for k in range(4)]
    return [m.astype(np.complex64) for m in masks], tts, gradient


class BioEdge(Pyramid):
    def __init__(self, nSubap: float, telescope, modulation: float, lightRatio: float,
                 postProcessing: str = 'slopesMaps', n_pix_separation: float = 2., n_pix_edge: float = None,
                 calibModulation: float = 50., extraModulationFactor: int = 0, binning: int = 1,
                 nTheta_user_defined: int = None, userValidSignal: bool = None, grey_width: float = 0,
                 grey_length: bool = False, delta_theta: float = 0., user_modulation_path: list = None,
                 quadrants_numbering: list = [0, 1, 2, 3], max_batch_bytes: float = 8e9):
        self.gpu_available = False
        self.nJobs = 1
        self.precision = np.float32
        self.precision_complex = np.complex64
        self.telescope = telescope
        self.device = telescope.device
        if (self.telescope.resolution / nSubap) % 2 != 0 and self.telescope.resolution / nSubap != 1:
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
        self.psfCentering = True
        self.binning = binning
        self.quadrants_numbering = list(quadrants_numbering)
        self.user_modulation_path = user_modulation_path
        self.modulation_length = 2 * np.pi
        self.old_mask = False
        self._weight_vector = None
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
        self.grey_width = grey_width
        self.grey_length = grey_length
        n = self.telescope.resolution
        self.resolution = int((self.nSubap * 2 + self.n_pix_separation + self.n_pix_edge * 2) * n / self.nSubap)
        self.zeroPaddingFactor = self.resolution / n
        self.zeroPadding = (self.resolution - n) // 2
        self.tag = 'pyramid'
        self.cam = Detector(round(nSubap * self.zeroPaddingFactor), device=self.device)
        self.focal_plane_camera = Detector(int((modulation * 4 + 12) * self.zeroPaddingFactor),
                                           psf_sampling=self.zeroPaddingFactor, device=self.device)
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
        lin = np.linspace(-np.pi, np.pi, n) * (n / self.telescope.initial_resolution)
        self._tip_lin = lin
        self._tilt_lin = lin
        cam = self.cam.resolution
        if self.resolution % cam:
            raise OopaoError('detector size %d does not divide N=%d' % (cam, self.resolution))
        # the four masks run one after another on one stream: one shared workspace
        ws = Workspace(self.device)
        self.engines = [PWFSEngine(n, self.resolution, cam, self.device, workspace=ws) for _ in range(4)]
        self.engine = self.engines[0]
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

    # ------------------------------------------------------------------ masks
    def mask_computation(self, sx=[0] * 4, sy=[0] * 4):
        """BioEdge.mask_computation (BioEdge.py:331-390)."""
        print('Bio-Edge Mask initialization...')
        masks, tts, grad = bioedge_mask_set(self.resolution, self.telescope.resolution,
                                            self.telescope.initial_resolution, self.nSubap, self.zeroPaddingFactor,
                                            self.grey_width, self.grey_length, sx, sy, self.quadrants_numbering)
        if grad is not None:
            self.gray_gradient = grad
        self.mask = masks
        self.mask_TT = tts
        self.initial_mask = [np.copy(m) for m in masks]
        zero = np.zeros((self.resolution, self.resolution))
        for e, m in zip(self.engines, masks):
            e.set_mask(zero, mask=m, general=True)
        print('Done!')

    def apply_shift_wfs(self, sx=None, sy=None, mis_reg=None, units='pixels'):
        """BioEdge.apply_shift_wfs (BioEdge.py:394-433)."""
        sx = 0 if sx is None else sx
        sy = 0 if sy is None else sy
        if mis_reg is not None:
            sx = [mis_reg.dX_1, mis_reg.dX_2, mis_reg.dX_3, mis_reg.dX_4]
            sy = [mis_reg.dY_1, mis_reg.dY_2, mis_reg.dY_3, mis_reg.dY_4]
        f = 2
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
        if np.max(np.abs(shift_x) / factor) > self.n_pix_edge or np.max(np.abs(shift_y) / factor) > self.n_pix_edge:
            warning('The Bi-O Edge pupils have been shifted outside of the detector!' +
                    'Wrapping of the signal is currently occuring!!')
        self.mask_computation(sx=sx, sy=sy)
        self.sx = np.asarray(shift_x) / factor
        self.sy = np.asarray(shift_y) / factor
        self.slopesUnits = 1
        self.referenceSignal = 0
        self._referenceSignal_2D = None
        self.wfs_calibration()

    @property
    def spatialFilter(self):
        return self._spatialFilter

    @spatialFilter.setter
    def spatialFilter(self, val):
        if val is not None:
            raise OopaoError('spatial filters are not supported for the Bi-O Edge on the GPU')
        self._spatialFilter = None

    # ------------------------------------------------------------ modulation
    @property
    def modulation(self):
        return self._modulation

    @modulation.setter
    def modulation(self, val):
        """BioEdge.modulation setter (BioEdge.py:967-1009), on the four engines."""
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
        for e in self.engines:
            e.set_modulation(path, self._tip_lin, self._tilt_lin, True)
        if self.isCalibrated:
            print('Updating the reference slopes and Wavelength Calibration for the new modulation')
            self.slopesUnits = 1
            self.referenceSignal = 0
            self._referenceSignal_2D = None
            self.wfs_calibration()
            print('Done!')

    # ---------------------------------------------------------- quadrants
    def _quadrant_origins(self):
        """grabQuadrant (BioEdge.py:840-853): w x w blocks around c = round(cam / binning / 2), w = c."""
        c = int(np.round((self.cam.resolution / self.binning) / 2))
        w = c
        rows = [c - w, c - w, c, c]
        cols = [c - w, c, c - w, c]
        return rows + cols, w

    def _setup_slopes(self):
        quad, w = self._quadrant_origins()
        self._w = w
        self.slopes_engine = SlopesEngine(self.device, (2 * self.cam.resolution) // self.binning, w, quad)

    def grabFullQuadrant(self, n, cameraFrame=None):
        """BioEdge.grabFullQuadrant (BioEdge.py:855-870)."""
        frame = self.cam.frame if cameraFrame is None else np.asarray(cameraFrame)
        n_tot = frame.shape[0]
        if n == 4:
            return frame[:n_tot // 2, :n_tot // 2]
        if n == 3:
            return frame[:n_tot // 2, -n_tot // 2:]
        if n == 1:
            return frame[-n_tot // 2:, -n_tot // 2:]
        if n == 2:
            return frame[-n_tot // 2:, :n_tot // 2]

    def initialization(self):
        """BioEdge.initialization (BioEdge.py:445-490)."""
        self.src ** self.telescope
        if self.userValidSignal is None:
            if self.lightRatio == 0:
                self.cam.frame = np.ones([self.cam.resolution * 2, self.cam.resolution * 2])
            else:
                print('The valid pixel are selected on flux considerations')
                self.modulation = self.calibModulation
                self.wfs_measure()
            self.initFrame = self.cam.frame
            if self.postProcessing[:10] == 'slopesMaps':
                self.I4Q = sum(self.grabQuadrant(k) for k in (1, 2, 3, 4))
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

    # ------------------------------------------------------------ measurement
    def _amp(self):
        """sqrt(0.5) * sqrt(src.intensity / nTheta) (BioEdge.py:514-520, 567)."""
        return (np.sqrt(0.5) * torch.sqrt(self.src._intensity / self.nTheta)).contiguous()

    def slope_bands(self):
        """The tiled four-mask frame is always computed whole."""
        return None

    def frames_dev(self, opd, weighted=False, bands=None, amp=None, scale=None):
        """Detector frames [B, 2 cam, 2 cam]: the four masks' binned images tiled 2 x 2 (``bands`` must
        be None: the Bi-O Edge frame is always whole)."""
        if bands is not None:
            raise OopaoError('column bands are not supported by the Bi-O Edge sensor')
        amp = self._amp() if amp is None else amp
        scale = 2 * np.pi / self.src.wavelength if scale is None else scale
        B = opd.shape[0]
        cam = self.cam.resolution
        out = torch.empty((B, 2 * cam, 2 * cam), dtype=torch.float32, device=self.device)
        bs = self._batch_size(B)
        while True:
            try:
                for j in range(0, B, bs):
                    o = opd[j:j + bs].contiguous()
                    for k, e in enumerate(self.engines):
                        r, c = divmod(k, 2)
                        out[j:j + bs, r * cam:(r + 1) * cam, c * cam:(c + 1) * cam] = e.frames(o, scale, amp,
                                                                                              weighted=weighted)
                return out
            except WorkspaceAllocationError as err:
                # as Pyramid.frames_dev: fewer screens per call, remembered in max_batch_bytes
                if bs == 1:
                    raise
                bs = max(1, min(bs, B) // 2)
                self.max_batch_bytes = bs * max(1, self.engine.bytes_per_screen())
                self.engine.workspace.release()      # shared by the four engines
                torch.cuda.empty_cache()
                warning('%s; measuring %d screens at a time' % (err.string, bs))

    def _raw_dev(self, args, weighted):
        opd, wavelength, amp = args
        N = self.resolution
        raw = torch.empty((2 * N, 2 * N), dtype=torch.float32, device=self.device)
        for k, e in enumerate(self.engines):
            r, c = divmod(k, 2)
            raw[r * N:(r + 1) * N, c * N:(c + 1) * N] = e.raw(opd[-1:].contiguous(), 2 * np.pi / wavelength, amp,
                                                              weighted=weighted)[0]
        return raw

    def focal_intensity_dev(self, window):
        """Focal-plane camera: the reference stores each point's unnormalised focal field once per mask
        (BioEdge.py:435-443, 1036-1045), i.e. 4 N^2 times the normalised intensity."""
        return super().focal_intensity_dev(window) * (4.0 * self.resolution ** 2)

    def __repr__(self):
        self.properties()
        n_char = len(max(self.prop.values(), key=len))
        return (f'\n{" Bi-O Edge WFS ":-^{n_char}}\n' + ''.join(v + '\n' for v in self.prop.values())
                + f'{"":-^{n_char}}\n')
