"""numpy restatement of OOPAO's modulated Pyramid WFS forward model (CPU oracle).

Test infrastructure only (see ``oracle/__init__.py``).

Two arithmetic modes are restated:

* ``'fp64'`` -- the reference as shipped (``OOPAO/__init__.py:55-60`` forces
  precision 64): complex128 field and FFTs, but the pyramid mask is stored as
  complex64 (``Pyramid.py:388``) and the modulation tip/tilt buffers as float32
  (``Pyramid.py:1133-1139``).
* ``'fp32'`` -- the reference with ``precision_oopao.npy`` = 32: complex64
  field, complex64 casts after each FFT (``Pyramid.py:629-631``), float32
  intensities summed in float32.

Pins (tests/test_oracle_pwfs.py): the mask and tables against their closed
forms evaluated along a different arithmetic route; the propagation against a
brute-force DFT on a tiny grid; Parseval flux conservation; zero slopes at the
calibration wavefront; flux-scale independence of the slopes; the sign of the
slope response to a small tip; the quadrant geometry against a hand count.
"""
import numpy as np
import scipy.fft

# Source.py:342-424 -- [wavelength (m), bandwidth (m), zero point (photons/m^2/s)]
_CST = 368.0
PHOTOMETRY = {
    'U': (0.360e-6, 0.070e-6, 1.96e12 / _CST),
    'B': (0.440e-6, 0.100e-6, 5.38e12 / _CST),
    'V0': (0.500e-6, 0.090e-6, 3.64e12 / _CST),
    'V': (0.550e-6, 0.090e-6, 3.31e12 / _CST),
    'R': (0.640e-6, 0.150e-6, 4.01e12 / _CST),
    'R2': (0.650e-6, 0.300e-6, 7.9e12 / _CST),
    'R3': (0.600e-6, 0.300e-6, 8.56e12 / _CST),
    'R4': (0.670e-6, 0.300e-6, 7.66e12 / _CST),
    'I': (0.790e-6, 0.150e-6, 2.69e12 / _CST),
    'I1': (0.700e-6, 0.033e-6, 0.67e12 / _CST),
    'I2': (0.750e-6, 0.033e-6, 0.62e12 / _CST),
    'I3': (0.800e-6, 0.033e-6, 0.58e12 / _CST),
    'I4': (0.700e-6, 0.100e-6, 2.02e12 / _CST),
    'I5': (0.850e-6, 0.100e-6, 1.67e12 / _CST),
    'I6': (1.000e-6, 0.100e-6, 1.42e12 / _CST),
    'I7': (0.850e-6, 0.300e-6, 5.00e12 / _CST),
    'I8': (0.750e-6, 0.100e-6, 1.89e12 / _CST),
    'I9': (0.850e-6, 0.300e-6, 5.00e12 / _CST),
    'I10': (0.900e-6, 0.300e-6, 4.72e12 / _CST),
    'J': (1.215e-6, 0.260e-6, 1.90e12 / _CST),
    'J2': (1.550e-6, 0.260e-6, 1.49e12 / _CST),
    'H': (1.654e-6, 0.290e-6, 1.05e12 / _CST),
    'Kp': (2.1245e-6, 0.351e-6, 0.62e12 / _CST),
    'Ks': (2.157e-6, 0.320e-6, 0.55e12 / _CST),
    'K': (2.179e-6, 0.410e-6, 0.70e12 / _CST),
    'K0': (2.000e-6, 0.410e-6, 0.76e12 / _CST),
    'K1': (2.400e-6, 0.410e-6, 0.64e12 / _CST),
    'L': (3.547e-6, 0.570e-6, 2.5e11 / _CST),
    'M': (4.769e-6, 0.450e-6, 8.4e10 / _CST),
    'Na': (0.589e-6, 0.0, 3.3e12 / _CST),
    'EOS': (1.064e-6, 0.0, 3.3e12 / _CST),
    'IR1310': (1.310e-6, 0.0, 2e12 / _CST),
}


def n_photon(band, magnitude):
    """Photons / m^2 / s of a source (Source.py:125)."""
    return PHOTOMETRY[band][2] * 10 ** (-0.4 * magnitude)


def telescope_pupil(resolution, central_obstruction=0.0):
    """Circular pupil with central obstruction (Telescope.py:228-235)."""
    D = resolution + 1
    x = np.linspace(-resolution / 2, resolution / 2, resolution)
    xx, yy = np.meshgrid(x, x)
    circle = xx ** 2 + yy ** 2
    obs = circle >= (central_obstruction * D / 2) ** 2
    return (circle < (D / 2) ** 2) * obs


def flux_map(pupil, nphoton, sampling_time, diameter):
    """Photons per pixel per frame (Telescope.py:223 with pupilReflectivity = pupil)."""
    n = pupil.shape[0]
    return pupil.astype(np.float64) * nphoton * sampling_time * (diameter / n) ** 2


class Geometry:
    """Pyramid geometry (Pyramid.py:216-217, 260-289, 292-296)."""

    def __init__(self, n, nSubap, n_pix_separation=2.0, n_pix_edge=None, calibModulation=50.0,
                 binning=1, fractional=False):
        # fractional: a non-integer n / nSubap (the reference raises; the detector then integrates by area)
        if (n / nSubap) % 2 != 0 and n / nSubap != 1 and not fractional:
            raise ValueError('The resolution should be an even number and be a multiple of 2**i where i>=2')
        self.n = n
        self.nSubap = nSubap
        self.sep = n_pix_separation
        self.edge = n_pix_separation // 2 if n_pix_edge is None else n_pix_edge
        full = (nSubap * 2 + self.sep + self.edge * 2) * n / nSubap
        self.N = int(round(full)) if fractional else int(full)
        self.zeroPaddingFactor = self.N / n
        self.center = self.N // 2
        self.cam = round(nSubap * self.zeroPaddingFactor)
        self.binning = binning
        self.calibModulation = n / 2 - 1 if calibModulation >= n / 2 else calibModulation


def tip_tilt(n, n_initial=None):
    """Unit tip/tilt grids for the modulation (Pyramid.py:334-339)."""
    n_initial = n if n_initial is None else n_initial
    tip, tilt = np.meshgrid(np.linspace(-np.pi, np.pi, n), np.linspace(-np.pi, np.pi, n))
    return tip * (n / n_initial), tilt * (n / n_initial)


def modulation_path(modulation, nTheta_user_defined=None, extraModulationFactor=0,
                    delta_theta=0.0, delta_Tip=0.0, delta_Tilt=0.0, user_modulation_path=None):
    """Modulation points in lambda/D (Pyramid.py:1117-1132). Returns [T, 2]."""
    if user_modulation_path is not None:
        return np.asarray(user_modulation_path, dtype=np.float64)
    perimeter = 2 * np.pi * modulation
    if nTheta_user_defined is None:
        nTheta = 4 * int(extraModulationFactor + np.ceil(perimeter / 4))
    else:
        nTheta = nTheta_user_defined
    theta = np.linspace(0 + delta_theta, 2 * np.pi + delta_theta, nTheta, endpoint=False)
    path = np.empty((nTheta, 2))
    for i in range(nTheta):
        path[i, 0] = modulation * np.cos(theta[i]) + delta_Tip
        path[i, 1] = modulation * np.sin(theta[i]) + delta_Tilt
    return path


def modulation_buffer(path, tip, tilt):
    """float32 tip/tilt phase per modulation point (Pyramid.py:1133-1139). [T, n, n]."""
    T = path.shape[0]
    n = tip.shape[0]
    buf = np.zeros([T, n, n], dtype=np.float32)
    for i in range(T):
        buf[i] = path[i, 0] * tip + path[i, 1] * tilt
    return buf


def phasor(N):
    """PSF-centring phasor exp(-i pi (N+1)/N (x+y)) (Pyramid.py:342-344)."""
    xx, yy = np.meshgrid(np.linspace(0, N - 1, N), np.linspace(0, N - 1, N))
    return np.exp(-(1j * np.pi * (N + 1) / N) * (xx + yy))


def phase_mask(geom, psfCentering=True, rooftop=0.0, theta_rotation=0.0, sx=(0, 0, 0, 0),
               sy=(0, 0, 0, 0)):
    """Pyramid phase mask m (radians) and complex64 exp(i m) (Pyramid.py:453-480, 388)."""
    rooftop_pixels = rooftop / np.sqrt(2)
    n_tot = geom.N
    norma = geom.n / geom.nSubap
    lim = np.pi
    x = np.linspace(-(1 - psfCentering * 1 / n_tot) * lim, (1 - psfCentering * 1 / n_tot) * lim,
                    n_tot, endpoint=psfCentering)
    x_, y_ = np.meshgrid(x, x)
    xr = x_ * np.cos(theta_rotation) - y_ * np.sin(theta_rotation)
    yr = y_ * np.cos(theta_rotation) + x_ * np.sin(theta_rotation)
    r = (geom.nSubap + geom.sep) / 2
    P1 = xr * r + x_ * sx[0] + yr * r - y_ * sy[0] + rooftop_pixels
    P2 = -xr * r + x_ * sx[1] + yr * r - y_ * sy[1]
    P3 = -xr * r + x_ * sx[2] - yr * r - y_ * sy[2] + rooftop_pixels
    P4 = xr * r + x_ * sx[3] - yr * r - y_ * sy[3]
    F = np.max(np.stack([P1, P2, P3, P4]) * norma, axis=0)
    m = -F
    return m, np.complex64(np.exp(1j * m))


def phase_mask_old(geom, psfCentering=True, sx=(0, 0, 0, 0), sy=(0, 0, 0, 0)):
    """The previous pyramid mask (Pyramid.py:482-539): a zero-mean tip/tilt ramp per quadrant.

    Restated per pixel from the closed forms of the centred ramps: with h = N/2,
    linspace(-pi, pi, h, endpoint=False) minus its mean is pi (2j - h + 1) / h;
    linspace(-pi, pi, h + 1) is already zero-mean, pi (2j - h) / h; and
    linspace(-l, l, h - 1, endpoint=False), l = pi - 2 pi / N, minus its mean is
    l (2j - h + 2) / (h - 1).
    """
    N = geom.N
    h = N // 2
    L = geom.nSubap + geom.sep
    norma = (geom.n / geom.nSubap) / 4
    m = np.zeros((N, N))

    def ramp(kind, j):
        j = np.asarray(j, np.float64)
        if kind == 'h':
            return np.pi * (2 * j - h + 1) / h
        if kind == 'p':
            return np.pi * (2 * j - h) / h
        lm = np.pi - 2 * np.pi / N
        return lm * (2 * j - h + 2) / (h - 1)

    def fill(r0, r1, c0, c1, kx, ky, a, b):
        rows = np.arange(r1 - r0)
        cols = np.arange(c1 - c0)
        tilt = ramp(ky, rows)[:, None]
        tip = ramp(kx, cols)[None, :]
        m[r0:r1, c0:c1] = (a * tip + b * tilt) * norma

    if psfCentering:
        fill(0, h, 0, h, 'h', 'h', L + sx[0], L - sy[0])
        fill(0, h, N - h, N, 'h', 'h', -(L - sx[1]), L - sy[1])
        fill(N - h, N, N - h, N, 'h', 'h', -(L - sx[2]), -(L + sy[2]))
        fill(N - h, N, 0, h, 'h', 'h', L + sx[3], -(L + sy[3]))
    else:
        fill(0, h + 1, 0, h + 1, 'p', 'p', L + sx[0], L - sy[0])
        fill(0, h + 1, h + 1, N, 'm', 'p', -(L - sx[1]), L - sy[1])
        fill(h + 1, N, h + 1, N, 'm', 'm', -(L - sx[2]), -(L + sy[2]))
        fill(h + 1, N, 0, h + 1, 'p', 'm', L + sx[3], -(L + sy[3]))
    m = -m
    return m, np.complex64(np.exp(1j * m))


def pyramid_transform(amp, phase_total, mask, phasor_nn, geom, precision='fp64', psfCentering=True,
                      spatial_filter=None):
    """One modulation point: field -> pad -> FFT -> mask -> IFFT -> |E|^2 (Pyramid.py:612-646).

    ``phase_total`` is src.phase + TT_i (already summed in the caller's precision).
    Returns (intensity [N,N], focal-plane field / N [N,N]).
    """
    n, N, c = geom.n, geom.N, geom.center
    cdt = np.complex128 if precision == 'fp64' else np.complex64
    support = np.zeros((N, N), dtype=cdt)
    em = amp * np.exp(1j * phase_total)
    support[c - n // 2:c + n // 2, c - n // 2:c + n // 2] = em
    if psfCentering:
        ft = scipy.fft.fft2(support * phasor_nn).astype(cdt)
    elif spatial_filter is not None:
        # Pyramid.py:633-634: the filter multiplies the shifted spectrum (and is also in the mask)
        ft = np.fft.fftshift(scipy.fft.fft2(support)) * spatial_filter
    else:
        ft = np.fft.fftshift(scipy.fft.fft2(support)).astype(cdt)
    pw = np.fft.ifft2(ft * mask).astype(cdt)
    return np.abs(pw) ** 2, ft / ft.shape[0]


def set_binning(array, binning_factor):
    """Sum-binning (tools/tools.py:464-483, mode='sum')."""
    if array.shape[0] % binning_factor != 0:
        raise ValueError('Binning factor %d not compatible with the array size' % binning_factor)
    new = [int(np.round(array.shape[0] / binning_factor)), int(np.round(array.shape[1] / binning_factor))]
    shape = (new[0], array.shape[0] // new[0], new[1], array.shape[1] // new[1])
    return array.reshape(shape).sum(-1).sum(1)


def bin_area(array, cam):
    """Detector integration by area onto cam x cam pixels: detector pixel R covers [R s, (R + 1) s) with
    s = N / cam and takes each raw pixel in proportion to the overlap; the sum binning of set_binning
    when cam divides N (tools/tools.py:464-483), flux-conserving for any s."""
    a = np.asarray(array, dtype=np.float64)
    N = a.shape[0]
    s = N / cam
    W = np.zeros((cam, N))
    for R in range(cam):
        lo, hi = R * s, (R + 1) * s
        for Y in range(int(np.floor(lo)), min(N, int(np.ceil(hi)))):
            W[R, Y] = min(Y + 1, hi) - max(Y, lo)
    return W @ a @ W.T


def grab_quadrant(frame, k, geom):
    """Quadrant k of the detector frame (Pyramid.py:946-961)."""
    b = geom.binning
    e = int(np.round((geom.sep / geom.nSubap) * geom.n / (geom.n / geom.nSubap) / 2 / b))
    c = int(np.round((geom.cam / b) / 2))
    w = int(np.ceil(geom.nSubap / b))
    if k == 3:
        return frame[e + c:e + c + w, e + c:e + c + w]
    if k == 4:
        return frame[e + c:e + c + w, -e + c - w:-e + c]
    if k == 1:
        return frame[-e + c - w:-e + c, -e + c - w:-e + c]
    if k == 2:
        return frame[-e + c - w:-e + c, e + c:e + c + w]
    raise ValueError(k)


def slopes_maps(frame, validI4Q, ref2d, geom, slopesUnits=1.0):
    """'slopesMaps' post-processing (Pyramid.py:838-854). Returns (signal_2D, signal)."""
    I = [grab_quadrant(frame, k, geom) * validI4Q for k in (1, 2, 3, 4)]
    I1, I2, I3, I4 = I
    I4Q = I1 + I2 + I3 + I4
    norma = np.mean(I4Q[validI4Q])
    Sx = I1 - I2 + I4 - I3
    Sy = I1 - I4 + I2 - I3
    sm = (np.concatenate((Sx, Sy) / norma) - ref2d) * slopesUnits
    valid_signal = np.concatenate((validI4Q, validI4Q))
    return sm, sm[np.where(valid_signal == 1)]


class PyramidOracle:
    """The reference Pyramid's numerical pipeline restated on numpy.

    Construction follows Pyramid.__init__ (Pyramid.py:213-366): valid pixels
    selected on a flat wavefront at calibModulation (initialization, 541-589),
    then the reference slopes at the working modulation (wfs_calibration,
    591-610). ``n_jobs`` > 1 dispatches the modulation points over a thread pool
    exactly like the reference's joblib loop (Pyramid.py:743-748).
    """

    def __init__(self, n, nSubap, modulation, lightRatio, pupil, amp_flux, wavelength,
                 n_pix_separation=2.0, n_pix_edge=None, calibModulation=50.0,
                 nTheta_user_defined=None, psfCentering=True, precision='fp64', n_jobs=1,
                 rooftop=0.0, theta_rotation=0.0, delta_theta=0.0, user_modulation_path=None,
                 validI4Q=None, sx=(0, 0, 0, 0), sy=(0, 0, 0, 0), spatial_filter=None, weight_vector=None,
                 old_mask=False, fractional=False):
        self.geom = Geometry(n, nSubap, n_pix_separation, n_pix_edge, calibModulation, fractional=fractional)
        self.precision = precision
        self.psfCentering = psfCentering
        self.pupil = pupil
        self.flux = amp_flux            # src.intensity (photons / px / frame)
        self.wavelength = wavelength
        self.n_jobs = n_jobs
        self.nTheta_user_defined = nTheta_user_defined
        self.delta_theta = delta_theta
        self.user_modulation_path = user_modulation_path
        self.tip, self.tilt = tip_tilt(n)
        self.phasor = phasor(self.geom.N)
        if old_mask:
            self.m, self.mask = phase_mask_old(self.geom, psfCentering, sx, sy)
        else:
            self.m, self.mask = phase_mask(self.geom, psfCentering, rooftop, theta_rotation, sx, sy)
        self.spatial_filter = spatial_filter
        if spatial_filter is not None:
            self.mask = self.mask * spatial_filter     # spatialFilter setter (Pyramid.py:1072)
        self.weight_vector = weight_vector
        self.lightRatio = lightRatio
        self._pool = None
        if validI4Q is not None:
            # timing use: take the valid pixels as given, reference slopes zero
            self.validI4Q = np.asarray(validI4Q, bool)
            self.validSignal = np.concatenate((self.validI4Q, self.validI4Q))
            self.nSignal = int(self.validSignal.sum())
            self.set_modulation(modulation)
            self.ref2d = 0
            return
        # initialization(): valid pixels at the calibration modulation
        self.set_modulation(self.geom.calibModulation)
        flat = np.zeros((n, n))
        frame = self.frame(flat)
        I4Q = sum(grab_quadrant(frame, k, self.geom) for k in (1, 2, 3, 4))
        self.I4Q = I4Q
        self.validI4Q = I4Q >= lightRatio * I4Q.max()
        self.validSignal = np.concatenate((self.validI4Q, self.validI4Q))
        self.nSignal = int(self.validSignal.sum())
        # modulation = modulation; wfs_calibration()
        self.set_modulation(modulation)
        self.ref2d = 0
        self.ref2d, self.ref = slopes_maps(self.frame(flat), self.validI4Q, 0, self.geom)

    def set_modulation(self, modulation):
        """Modulation setter (Pyramid.py:1107-1145)."""
        if modulation >= self.geom.n // 2:
            raise ValueError('Error the modulation radius is too large for this resolution!')
        self.modulation = modulation
        if modulation != 0 or self.user_modulation_path is not None:
            self.path = modulation_path(modulation, self.nTheta_user_defined, 0, self.delta_theta,
                                        user_modulation_path=self.user_modulation_path)
            self.tt = modulation_buffer(self.path, self.tip, self.tilt)
        else:
            self.path = np.zeros((1, 2))
            self.tt = None
        self.nTheta = self.path.shape[0]

    def _amp(self):
        dt = np.float64 if self.precision == 'fp64' else np.float32
        return np.sqrt(self.flux.astype(dt) / self.nTheta)

    def raw(self, phase):
        """Modulated raw frame Sum_T |E_T|^2 for src.phase = ``phase`` [n,n] (Pyramid.py:712-757)."""
        amp = self._amp()
        rdt = np.float64 if self.precision == 'fp64' else np.float32
        ph = np.asarray(phase, dtype=rdt)
        sf = self.spatial_filter
        if self.tt is None:
            return pyramid_transform(amp, ph, self.mask, self.phasor, self.geom, self.precision,
                                     self.psfCentering, sf)[0]

        def one(i):
            return pyramid_transform(amp, ph + self.tt[i], self.mask, self.phasor, self.geom,
                                     self.precision, self.psfCentering, sf)[0]
        if self.n_jobs > 1:
            from concurrent.futures import ThreadPoolExecutor
            if self._pool is None:
                self._pool = ThreadPoolExecutor(self.n_jobs)
            maps = list(self._pool.map(one, range(self.nTheta)))
        else:
            maps = [one(i) for i in range(self.nTheta)]
        if self.weight_vector is not None:
            # Pyramid.py:753-755: diag(w) @ maps, summed, divided by nTheta
            w = np.asarray(self.weight_vector, np.float64)
            return np.tensordot(w, np.asarray(maps), axes=1) / self.nTheta
        return np.sum(np.asarray(maps), axis=0)

    def focal(self, phase, resolution):
        """Modulation focal-plane camera frame (Pyramid.py:642, 1183-1190): sum over the points of
        |fft2(support phasor) / N|^2, centre-cropped to ``resolution`` (whole field if larger than N)."""
        amp = self._amp()
        rdt = np.float64 if self.precision == 'fp64' else np.float32
        ph = np.asarray(phase, dtype=rdt)
        tts = [0] if self.tt is None else [self.tt[i] for i in range(self.nTheta)]
        inten = 0
        for tt in tts:
            ft = pyramid_transform(amp, ph + tt, self.mask, self.phasor, self.geom, self.precision,
                                   self.psfCentering, self.spatial_filter)[1]
            inten = inten + np.abs(ft) ** 2
        N = inten.shape[0]
        if resolution > N:
            return inten
        return inten[N // 2 - resolution // 2:N // 2 + resolution // 2,
                     N // 2 - resolution // 2:N // 2 + resolution // 2]

    def frame(self, phase):
        """Detector frame: sum-binning (by area for a non-integer binning), QE = gain = 1, float64
        (Pyramid.py:1193-1203, Detector.py:273-354)."""
        raw = self.raw(phase)
        if self.geom.N % self.geom.cam:
            return bin_area(raw, self.geom.cam)
        return set_binning(raw, self.geom.N / self.geom.cam).astype(np.float64)

    def measure(self, phase):
        """(signal_2D, signal) for one phase screen (Pyramid.py:666-757, 835-854)."""
        return slopes_maps(self.frame(phase), self.validI4Q, self.ref2d, self.geom)

    def measure_opd(self, opd):
        """Signal for an OPD in metres: src.phase = OPD*pupil*2*pi/lambda (Source.py:294-296)."""
        return self.measure(np.asarray(opd) * self.pupil * 2 * np.pi / self.wavelength)
