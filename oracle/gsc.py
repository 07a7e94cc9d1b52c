"""numpy restatement of OOPAO's gain-sensing camera (CPU oracle, complex128).

Test infrastructure only (see ``oracle/__init__.py``).

Follows GainSensingCamera.py:85-120 (calibration, compute_optical_gains) and 147-194 (padder,
fft3/ifft3, basis product, impulse response, sensitivity, optical gains). The centred transform
is written here as fftshift(fft2(fftshift(x))) / N with numpy's FFT.

Pins (tests/test_oracle_gsc.py): ``centred_fft2`` / ``centred_ifft2`` against the explicit
centred DFT matrices D[k, j] = exp(-+2 pi i (k - N/2)(j - N/2) / N) on small grids (an
independent route to the same transform, fixes the shift placement); ``basis_product`` against
|F(B)|^2 for a real basis and Parseval (sum over frequencies = sum |B|^2); ``impulse_response``
is identically zero for a flat mask (F(1) is a delta at the centre and the frame is real);
``GSCOracle`` gives og = 1 at the calibration frame and is invariant to the frame's scale.
"""
import numpy as np


def centred_fft2(x):
    """fftshift(fft2(fftshift(x))) / N over the first two axes."""
    N = x.shape[0]
    return np.fft.fftshift(np.fft.fft2(np.fft.fftshift(x, axes=(0, 1)), axes=(0, 1)), axes=(0, 1)) / N


def centred_ifft2(x):
    """fftshift(ifft2(fftshift(x))) * N over the first two axes."""
    N = x.shape[0]
    return np.fft.fftshift(np.fft.ifft2(np.fft.fftshift(x, axes=(0, 1)), axes=(0, 1)), axes=(0, 1)) * N


def pad_basis(basis, N):
    """Zero-pad [n, n, M] maps symmetrically to [N, N, M] (GainSensingCamera.py:139-145)."""
    n = basis.shape[0]
    if N <= n:
        raise ValueError('output size must exceed the input size')
    p = (N - n) // 2
    return np.pad(basis, ((p, p), (p, p), (0, 0)))


def basis_product(basis, N):
    """F(B_m) F'(B_m) for every mode of the padded basis, complex [N, N, M] (GainSensingCamera.py:155-171)."""
    B = pad_basis(np.asarray(basis, dtype=np.float64), N)
    return centred_fft2(B) * centred_ifft2(B)


def impulse_response(mask, frame):
    """2 Im(conj(F(mask)) F(mask frame)) (GainSensingCamera.py:179-180); frame already normalised."""
    return 2 * np.imag(np.conj(centred_fft2(mask)) * centred_fft2(mask * frame))


def sensitivity(ir1, ir2, bp):
    """(F(ir1) F'(ir2)).flatten() @ bp.reshape(N^2, M) (GainSensingCamera.py:183-188)."""
    w = (centred_fft2(ir1) * centred_ifft2(ir2)).reshape(-1)
    return w @ bp.reshape(-1, bp.shape[-1])


class GSCOracle:
    def __init__(self, mask, basis):
        self.mask = np.asarray(mask, dtype=np.complex128)
        self.basis = np.asarray(basis, dtype=np.float64)

    def calibration(self, frame):
        frame = np.asarray(frame, dtype=np.float64)
        frame = frame / frame.sum()
        self.bp = basis_product(self.basis, self.mask.shape[0])
        self.ir_calib = impulse_response(self.mask, frame)
        self.s_calib = sensitivity(self.ir_calib, self.ir_calib, self.bp)

    def optical_gains(self, frame):
        frame = np.asarray(frame, dtype=np.float64)
        frame = frame / frame.sum()
        self.ir_sky = impulse_response(self.mask, frame)
        s = sensitivity(self.ir_sky, self.ir_calib, self.bp)
        return np.real(s / self.s_calib)
