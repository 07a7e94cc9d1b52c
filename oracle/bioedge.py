"""numpy restatement of OOPAO's Bi-O Edge wavefront sensor (CPU oracle, float64).

Test infrastructure only (see ``oracle/__init__.py``).

Follows BioEdge.py: geometry (276-289), masks (mask_computation 331-390: knife-edge amplitude
A = sqrt(BW(x)) with the grey gradient, B = sqrt(1 - A^2), C = A^T, D = B^T, grey_length,
quadrants_numbering, tip/tilt shifts), the transform (bio_edge_transform 510-544: one focal field
per modulation point, four masked inverse transforms tiled 2 x 2 into a 2N x 2N frame, field
amplitude sqrt(0.5) sqrt(I / nTheta)), the detector binning (__mul__ 1033-1055), the quadrants
(grabQuadrant 840-853) and the post-processing (signalProcessing 729-820).

Pins (tests/test_oracle_bioedge.py): the four masks split the focal energy exactly
(A^2 + B^2 = C^2 + D^2 = 1), so the raw frame's total equals the pupil flux (Parseval); a
tip/tilt shift of sx pupil pixels on a mask rolls that mask's image by sx * n / nSubap raw
pixels exactly (integer case); with zero grey width the masks are complementary half planes.
"""
import numpy as np

from .pwfs import modulation_buffer, modulation_path, phasor, set_binning, tip_tilt


class BioEdgeGeometry:
    def __init__(self, n, nSubap, n_pix_separation=2.0, n_pix_edge=None):
        self.n, self.nSubap = n, nSubap
        self.sep = n_pix_separation
        self.edge = n_pix_separation // 2 if n_pix_edge is None else n_pix_edge
        self.N = int((nSubap * 2 + self.sep + self.edge * 2) * n / nSubap)
        self.zpf = self.N / n
        self.cam = int(round(nSubap * self.zpf))


def bioedge_masks(g, grey_width=0.0, grey_length=False, sx=(0, 0, 0, 0), sy=(0, 0, 0, 0),
                  quadrants_numbering=(0, 1, 2, 3), n_initial=None):
    """The four complex masks [4, N, N] (BioEdge.py:331-390)."""
    N, n = g.N, g.n
    n_initial = n if n_initial is None else n_initial
    BW = np.zeros(N)
    BW[:N // 2] = 1.0
    r = int(np.round(g.zpf * grey_width * n / n_initial))
    if grey_width != 0:
        grad = np.hstack([np.linspace(1, 0.5, r, endpoint=True), np.linspace(0.5, 0, r, endpoint=True)])
        BW[N // 2 - r:N // 2 + r] = grad
    X = np.tile(BW, (N, 1))                       # varies along columns (x)
    A = np.sqrt(X)
    if grey_length is not False:
        rl = int(np.round(g.zpf * grey_length))
        A[:N // 2 - rl, :N // 2] = 1.0
        A[N // 2 + rl:, :N // 2] = 1.0
        A[:N // 2 - rl, N // 2:] = 0.0
        A[N // 2 + rl:, N // 2:] = 0.0
    B = np.sqrt(1 - A ** 2)
    C, D = A.T.copy(), B.T.copy()
    lin = np.linspace(-np.pi, np.pi, N, endpoint=False)
    tip = np.tile(lin, (N, 1))
    tilt = tip.T
    norma = (g.nSubap * 2) * (n / g.nSubap) * (1 / (g.nSubap * 2))
    sx = [sx] * 4 if np.isscalar(sx) else list(sx)
    sy = [sy] * 4 if np.isscalar(sy) else list(sy)
    tt = [-sx[3 - k] * tip * norma + sy[3 - k] * tilt * norma for k in range(4)]   # A_TT uses index 3, ...
    amp = [A, B, C, D]
    amp = [amp[quadrants_numbering[k]] for k in range(4)]
    return np.stack([amp[k] * np.exp(1j * tt[k]) for k in range(4)])


class BioEdgeOracle:
    def __init__(self, n, nSubap, modulation, lightRatio, pupil, amp_flux, postProcessing='slopesMaps',
                 n_pix_separation=2.0, n_pix_edge=None, calibModulation=50.0, nTheta_user_defined=None,
                 grey_width=0.0, grey_length=False, sx=(0, 0, 0, 0), sy=(0, 0, 0, 0),
                 quadrants_numbering=(0, 1, 2, 3), incidence_norma=None):
        self.g = BioEdgeGeometry(n, nSubap, n_pix_separation, n_pix_edge)
        self.pupil = pupil
        self.flux = amp_flux
        self.pp = postProcessing
        self.nTheta_user_defined = nTheta_user_defined
        self.tip, self.tilt = tip_tilt(n)
        self.phasor = phasor(self.g.N)
        self.masks = bioedge_masks(self.g, grey_width, grey_length, sx, sy, quadrants_numbering)
        self.incidence_norma = incidence_norma      # nPhoton * samplingTime * (D / nSubap)^2
        calib = n / 2 - 1 if calibModulation >= n / 2 else calibModulation
        # initialization (BioEdge.py:445-490)
        flat = np.zeros((n, n))
        self.ref2d = 0
        if lightRatio == 0:
            self.init_frame = np.ones((2 * self.g.cam, 2 * self.g.cam))
            self.set_modulation(modulation)
        else:
            self.set_modulation(calib)
            self.init_frame = self.frame(flat)
        if self.pp[:10] == 'slopesMaps':
            I4Q = sum(self.quadrant(self.init_frame, k) for k in (1, 2, 3, 4))
            self.validI4Q = I4Q >= lightRatio * I4Q.max()
            self.validSignal = np.concatenate((self.validI4Q, self.validI4Q))
        else:
            self.validSignal = self.init_frame >= lightRatio * self.init_frame.max()
        # wfs_calibration (BioEdge.py:492-508) at the working modulation
        self.set_modulation(modulation)
        self.ref2d, _ = self.process(self.frame(flat))

    def set_modulation(self, modulation):
        self.modulation = modulation
        if modulation != 0:
            self.path = modulation_path(modulation, self.nTheta_user_defined)
            self.tt = modulation_buffer(self.path, self.tip, self.tilt)
            self.nTheta = self.path.shape[0]
        else:
            self.tt = None
            self.nTheta = 1

    def raw(self, phase):
        """2N x 2N raw frame summed over the modulation points (BioEdge.py:510-544, 609-654)."""
        g = self.g
        N, n = g.N, g.n
        amp = np.sqrt(0.5) * np.sqrt(self.flux / self.nTheta)
        c = N // 2
        out = np.zeros((2 * N, 2 * N))
        tts = [0.0] if self.tt is None else [self.tt[i].astype(np.float64) for i in range(self.nTheta)]
        for tt in tts:
            sup = np.zeros((N, N), complex)
            sup[c - n // 2:c + n // 2, c - n // 2:c + n // 2] = amp * np.exp(1j * (phase + tt))
            F = np.fft.fft2(sup * self.phasor)
            for k in range(4):
                i, j = divmod(k, 2)
                out[i * N:(i + 1) * N, j * N:(j + 1) * N] += np.abs(np.fft.ifft2(F * self.masks[k])) ** 2
        return out

    def frame(self, phase):
        return set_binning(self.raw(phase), self.g.N / self.g.cam)

    def focal(self, phase, resolution):
        """Focal-plane camera (BioEdge.py:1036-1045): the unnormalised focal field of each point is
        stored once per mask, so the frame is 4 sum_T |fft2(support phasor)|^2, centre-cropped."""
        g = self.g
        N, n = g.N, g.n
        amp = np.sqrt(0.5) * np.sqrt(self.flux / self.nTheta)
        c = N // 2
        tts = [0.0] if self.tt is None else [self.tt[i].astype(np.float64) for i in range(self.nTheta)]
        I = np.zeros((N, N))
        for tt in tts:
            sup = np.zeros((N, N), complex)
            sup[c - n // 2:c + n // 2, c - n // 2:c + n // 2] = amp * np.exp(1j * (phase + tt))
            I += 4 * np.abs(np.fft.fft2(sup * self.phasor)) ** 2
        if resolution > N:
            return I
        return I[N // 2 - resolution // 2:N // 2 + resolution // 2, N // 2 - resolution // 2:N // 2 + resolution // 2]

    def quadrant(self, frame, k):
        """grabQuadrant (BioEdge.py:840-853): the four w x w blocks around the point c = round(cam/2)."""
        c = int(np.round(self.g.cam / 2))
        w = c
        r0, c0 = {1: (c - w, c - w), 2: (c - w, c), 3: (c, c - w), 4: (c, c)}[k]
        return frame[r0:r0 + w, c0:c0 + w]

    def process(self, frame):
        """signalProcessing (BioEdge.py:729-820) for slopesMaps, slopesMaps_incidence_flux and the
        fullFrame variants; returns (signal_2D, signal)."""
        pp = self.pp
        if pp[:10] == 'slopesMaps':
            v = self.validI4Q
            I1, I2, I3, I4 = (self.quadrant(frame, k) * v for k in (1, 2, 3, 4))
            if pp == 'slopesMaps':
                norma = np.mean((I1 + I2 + I3 + I4)[v])
            elif pp == 'slopesMaps_incidence_flux':
                norma = self.incidence_norma
            else:
                norma = frame.mean()
            Sx = I1 - I2 + I4 - I3
            Sy = I1 - I4 + I2 - I3
            s2d = np.concatenate((Sx, Sy)) / norma - self.ref2d
            return s2d, s2d[self.validSignal]
        if pp == 'fullFrame_camera_flux':
            norma = frame.mean()
        elif pp == 'fullFrame_incidence_flux':
            norma = self.incidence_norma / 4
        elif pp == 'fullFrame_sum_flux':
            norma = frame.sum()
        else:
            norma = np.sum(frame[self.validSignal])
        s2d = frame / norma - self.ref2d
        return s2d, s2d[self.validSignal]

    def measure(self, phase):
        return self.process(self.frame(phase))
