"""GainSensingCamera (GainSensingCamera.py:13-199): optical gains of the modulated PWFS from the
modulation-camera frame with the convolutional model of Chambouleyron et al. 2021, on the GPU.

With F(x) = fftshift(fft2(fftshift(x))) / N and F'(x) = fftshift(ifft2(fftshift(x))) * N
(GainSensingCamera.py:147-152):

* calibration (95-106): basis product bp_m = F(B_m) F'(B_m) of every mode zero-padded to the
  mask size; IR_calib = 2 Im(conj(F(mask)) F(mask frame / sum frame)) (179-180); S_calib =
  (F(IR_calib) F'(IR_calib)) @ bp (183-188);
* compute_optical_gains (114-120): IR_sky of the new frame, S_sky = (F(IR_sky) F'(IR_calib)) @ bp,
  og = Re(S_sky / S_calib) (191-194).

For a real modal basis F'(B) = conj(F(B)), so bp = |F(B_m)|^2 is real; it is kept on the device
as float32 [N^2, nModes] (the reference holds complex128 [N, N, nModes]) and read once per
sensitivity. FFTs are complex64 (oopao_fft2), sums are accumulated in double. A complex basis
raises OopaoError.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._device import default_device, ptr, stream_ptr
from .tools import OopaoError


def _fft2(x, dir, scale, ws):
    N = x.shape[-1]
    out = torch.empty_like(x)
    _lib.call('oopao_fft2', ptr(x), ptr(out), N, 1, dir, 1, ctypes.c_float(scale), ptr(ws), ws.numel(),
              stream_ptr(x.device))
    return out


class GainSensingCamera:
    def __init__(self, mask: np.array, basis: np.array, n_jobs: int = 10) -> None:
        self.mask = mask
        self.basis = basis
        self.n_modes = self.basis.shape[-1]
        self.n_jobs = n_jobs
        self.calibration_ready = False
        self.detector_properties = None
        self.tag = 'GSC'
        self.device = default_device()
        self.modes_per_batch = 64

    # ------------------------------------------------------------ internals
    def _setup(self):
        m = self.mask.detach().cpu().numpy() if isinstance(self.mask, torch.Tensor) else np.asarray(self.mask)
        if m.ndim != 2 or m.shape[0] != m.shape[1] or m.shape[0] % 2:
            raise OopaoError('the GSC needs a square mask of even size, got %s' % (m.shape,))
        b = self.basis.detach().cpu().numpy() if isinstance(self.basis, torch.Tensor) else np.asarray(self.basis)
        if np.iscomplexobj(b):
            raise OopaoError('the GSC on the GPU takes a real modal basis [n, n, nModes]')
        N, n = m.shape[0], b.shape[0]
        if N <= n:
            raise OopaoError(f'Output size {N} should be higher than input size {n}')
        if (N - n) % 2:
            raise OopaoError('mask size minus basis size must be even (symmetric zero padding)')
        _lib.prepare_tables(N, self.device)
        if not _lib.lib().oopao_fft2_supported(N):
            raise OopaoError('no FFT plan for N=%d' % N)
        self.N, self.n = N, n
        dev = self.device
        self._mask = torch.from_numpy(np.ascontiguousarray(m.astype(np.complex64))).to(dev)
        nb = ctypes.c_size_t(0)
        mb = max(1, min(self.modes_per_batch, self.n_modes))
        _lib.call('oopao_gsc_workspace_bytes', N, self.n_modes, mb, ctypes.byref(nb))
        self._ws = torch.empty(nb.value, dtype=torch.uint8, device=dev)
        self._mb = mb
        self._fmask = _fft2(self._mask, -1, 1.0 / N, self._ws)
        modes = torch.from_numpy(np.ascontiguousarray(np.moveaxis(b, -1, 0), dtype=np.float32)).to(dev)
        self._bp = torch.empty((N * N, self.n_modes), dtype=torch.float32, device=dev)
        for m0 in range(0, self.n_modes, mb):
            k = min(mb, self.n_modes - m0)
            _lib.call('oopao_gsc_basis_product', ptr(modes), n, N, self.n_modes, m0, k, ptr(self._bp), ptr(self._ws),
                      self._ws.numel(), stream_ptr(dev))

    def _frame_dev(self, frame):
        f = frame if isinstance(frame, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(frame, np.float32))
        f = f.to(self.device, torch.float32).contiguous()
        if f.shape != (self.N, self.N):
            raise OopaoError('frame must be %dx%d (set focal_plane_camera.resolution = wfs.nRes)' % (self.N, self.N))
        return f

    def _impulse_response(self, frame):
        ir = torch.empty((self.N, self.N), dtype=torch.complex64, device=self.device)
        _lib.call('oopao_gsc_impulse_response', ptr(self._fmask), ptr(self._mask), ptr(self._frame_dev(frame)),
                  self.N, ptr(ir), ptr(self._ws), self._ws.numel(), stream_ptr(self.device))
        return ir

    def _sensitivity(self, ir1, v2):
        S = torch.empty(2 * self.n_modes, dtype=torch.float64, device=self.device)
        _lib.call('oopao_gsc_sensitivity', ptr(ir1), ptr(v2), ptr(self._bp), self.N, self.n_modes, ptr(S),
                  ptr(self._ws), self._ws.numel(), stream_ptr(self.device))
        return S.view(self.n_modes, 2)

    # ------------------------------------------------------------ reference API
    def calibration(self, frame) -> None:
        """GainSensingCamera.calibration (GainSensingCamera.py:85-100)."""
        print('\nGain Scheduling camera calibration:')
        self._setup()
        self._ir_calib = self._impulse_response(frame)
        self._v_calib = _fft2(self._ir_calib, 1, 1.0 / self.N, self._ws)     # F'(IR_calib)
        self._s_calib = self._sensitivity(self._ir_calib, self._v_calib)
        self.calibration_ready = True
        self.og = np.ones(self.n_modes)

    def reset_calibration(self) -> None:
        self.calibration_ready = False
        self.og = np.ones(self.n_modes)

    def compute_optical_gains(self, frame) -> None:
        """GainSensingCamera.compute_optical_gains (GainSensingCamera.py:108-120)."""
        if self.calibration_ready is False:
            raise OopaoError('Optical gains must be initialized first!')
        self._ir_sky = self._impulse_response(frame)
        s = self._sensitivity(self._ir_sky, self._v_calib)
        c = self._s_calib
        # Re(s / c) = (s.re c.re + s.im c.im) / |c|^2
        og = (s[:, 0] * c[:, 0] + s[:, 1] * c[:, 1]) / (c[:, 0] ** 2 + c[:, 1] ** 2)
        self.og = og.cpu().numpy()

    # ------------------------------------------------------------ host copies
    @property
    def basis_product(self):
        """|F(B_m)|^2 as [N, N, nModes] float64 (the reference's basis product, real for a real basis)."""
        return self._bp.double().cpu().numpy().reshape(self.N, self.N, self.n_modes)

    @property
    def IR_calib(self):
        return self._ir_calib.real.double().cpu().numpy()

    @property
    def IR_sky(self):
        return self._ir_sky.real.double().cpu().numpy()

    @property
    def sensi_calib(self):
        c = self._s_calib.cpu().numpy()
        return c[:, 0] + 1j * c[:, 1]

    def __repr__(self) -> str:
        prop = f'{"Calibrated":<16}|{str(self.calibration_ready):^7}\n'
        prop += f'{"Number of modes":<16}|{self.n_modes:^7}\n'
        n_char = len(max(prop.split('\n'), key=len))
        return f'{" GSC ":-^{n_char}}\n' + prop + f'{"":-^{n_char}}\n'
