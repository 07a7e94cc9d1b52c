"""Science path on the GPU: PSF of a source through the telescope and the OTF Strehl ratio.

Telescope.PropagateField (Telescope.py:374-423) zero-pads the pupil field
amp * exp(i phase) to N = zeroPaddingFactor * oversampling * n samples,
multiplies by the centring phasor exp(-i pi (N+1)/N (x+y)) and keeps the
centred img x img window of |fft2(.)/N|^2. That is the modulation focal-plane
camera of the general PWFS kernels with a single "modulation point" whose tip
and tilt tables are the phasor's row and column factors, so the PSF is
computed by oopao_pwfs_generic_focal (row transforms of the pupil rows, then
only the window's column transforms). strehlMeter (tools/tools.py:102-157)
compares the OTF of the image with the OTF of the diffraction-limited image; the
OTF magnitude |fft2(image)| comes from the same focal kernel (a unit phasor, the
whole image as window).
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from ._device import ptr, stream_ptr
from .tools import OopaoError


def psf_geometry(n, zeroPaddingFactor, img_resolution):
    """(oversampling, padded N, pad offset, image size) of Telescope.PropagateField (Telescope.py:375-393)."""
    over = 1
    if zeroPaddingFactor * over < 2:
        over = int(math.ceil(2.0 / zeroPaddingFactor))
    if img_resolution is None:
        img_resolution = zeroPaddingFactor * n
    if img_resolution > zeroPaddingFactor * n:
        raise OopaoError('Error: image has too many pixels for this pupil sampling. Try using a pupil mask with '
                         'more pixels')
    img_size = int(math.ceil(img_resolution * over))
    N = int(np.fix(zeroPaddingFactor * over * n))
    pad = int(math.ceil((N - n) / 2))
    N = n + 2 * pad
    if N % 2 or img_size % 2 or n % 2:
        raise OopaoError('PSF computation supports even pupil, padded and image sizes (n=%d, N=%d, image %d)'
                         % (n, N, img_size))
    return over, N, pad, img_size


class _FocalEngine:
    """Cached tables and workspace for |fft2(pad(amp exp(i scale opd)) tab_y tab_x) / N|^2 on a centred window
    (the general focal kernel with one point)."""

    def __init__(self, n, N, img, tab, device):
        self.n, self.N, self.img, self.device = n, N, img, device
        if not _lib.lib().oopao_pwfs_generic_supported(n, N, 1):
            raise OopaoError('padded size N=%d outside the supported range' % N)
        self.tab = torch.from_numpy(np.ascontiguousarray(tab.astype(np.complex64)[None, :])).to(device)
        _lib.prepare_tables(N, device)
        nb = ctypes.c_size_t(0)
        _lib.call('oopao_pwfs_generic_focal_bytes', n, N, 1, 1, 1, img, ctypes.byref(nb))
        self.nb = nb.value
        self.ws = torch.empty(self.nb, dtype=torch.uint8, device=device)
        self.zero = torch.zeros((n, n), dtype=torch.float32, device=device)

    def __call__(self, amp, opd, scale):
        out = torch.empty((1, self.img, self.img), dtype=torch.float32, device=self.device)
        opd = self.zero if opd is None else opd.contiguous()
        _lib.call('oopao_pwfs_generic_focal', ptr(opd), ctypes.c_float(scale), ptr(amp.contiguous()),
                  ptr(self.tab), ptr(self.tab), self.n, self.N, 1, 1, 1, self.img, ptr(self.ws),
                  ctypes.c_size_t(self.nb), ptr(out), stream_ptr(self.device))
        return out[0]


_ENGINES = {}


def _engine(key, make):
    e = _ENGINES.get(key)
    if e is None:
        e = _ENGINES[key] = make()
    return e


def psf_dev(amp, opd, scale, n, zeroPaddingFactor, img_resolution, device):
    """|fft2(pad(amp exp(i scale opd)) phasor) / N|^2 on the centred image window, float32 [img, img]."""
    over, N, pad, img = psf_geometry(n, zeroPaddingFactor, img_resolution)
    key = ('psf', str(device), n, N, img)

    def make():
        X = np.arange(n) + pad
        return _FocalEngine(n, N, img, np.exp(-1j * np.pi * (N + 1) / N * X), device)
    psf = _engine(key, make)(amp.float(), opd.float(), scale)
    if over != 1:
        r = img // over
        psf = psf.reshape(r, over, r, over).sum(dim=(1, 3))
    return psf


def otf_ratio_dev(img):
    """sum |OTF| / max |OTF| of an image, OTF = fft2(image), through the general focal kernel (0-d tensor)."""
    s = img.shape[0]
    e = _engine(('otf', str(img.device), s), lambda: _FocalEngine(s, s, s, np.ones(s), img.device))
    otf = torch.sqrt(e(img.float(), None, 0.0).double())
    return otf.sum() / otf.max()


def strehl_dev(psf, psf_ref, resolution, zeroPaddingFactor=2, as_tensor=False):
    """Strehl ratio [%] from the OTFs of an image and of the reference image (tools/tools.py:102-157);
    a 0-d device tensor with ``as_tensor`` (no host synchronisation)."""
    s = min(psf_ref.shape[0], psf.shape[0])

    def crop(a):
        c = a.shape[0] // 2
        return a[c - s // 2:c + s // 2, c - s // 2:c + s // 2].contiguous()
    sr = otf_ratio_dev(crop(psf)) / otf_ratio_dev(crop(psf_ref)) * 100
    return sr if as_tensor else float(sr)
