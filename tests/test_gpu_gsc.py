"""GPU: 2-D FFT (oopao_amd/csrc/fft2.hip) and the gain-sensing camera (csrc/gsc.hip,
oopao_amd/GainSensingCamera.py) against numpy and the float64 oracle (oracle/gsc.py).

Reference: GainSensingCamera.py:78-199, Detector.__mul__ (Detector.py:385-393).
Tolerances (complex64 FFTs, float32 basis product, double sums): FFT rel-L2 <= 1e-5 vs numpy
complex128; basis product and impulse response rel-L2 <= 1e-5; optical gains within 1e-4 of the
oracle's.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle.gsc import GSCOracle, basis_product
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('N,centred', [(64, True), (96, True), (80, False), (126, True), (128, False)])
def test_fft2_matches_numpy(N, centred):
    from oopao_amd import _lib
    from oopao_amd._device import ptr, stream_ptr
    print('fft2', N, centred, flush=True)
    rs = np.random.RandomState(N)
    B = 3
    x = (rs.randn(B, N, N) + 1j * rs.randn(B, N, N)).astype(np.complex64)
    xd = torch.from_numpy(x).cuda()
    nb = ctypes.c_size_t(0)
    _lib.call('oopao_fft2_bytes', N, B, ctypes.byref(nb))
    ws = torch.empty(nb.value, dtype=torch.uint8, device='cuda')
    sh = (lambda a: np.fft.fftshift(a, axes=(1, 2))) if centred else (lambda a: a)
    for d, ref in ((-1, sh(np.fft.fft2(sh(x.astype(np.complex128)), axes=(1, 2))) / N),
                   (1, sh(np.fft.ifft2(sh(x.astype(np.complex128)), axes=(1, 2))) * N)):
        out = torch.empty_like(xd)
        _lib.call('oopao_fft2', ptr(xd), ptr(out), N, B, d, int(centred), ctypes.c_float(1.0 / N), ptr(ws),
                  ws.numel(), stream_ptr(xd.device))
        err = rel_l2(out.cpu().numpy(), ref)
        print('dir', d, 'rel-L2', err)
        assert err < 1e-5


def test_gsc_against_oracle():
    from oopao_amd import Telescope, Source, DeformableMirror, Pyramid
    from oopao_amd.GainSensingCamera import GainSensingCamera
    print('gsc n=32 N=128', flush=True)
    tel = Telescope(32, 8, 1e-3)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    dm = DeformableMirror(tel, 8, 0.35, print_dm_properties=False)
    wfs = Pyramid(8, tel, 3, 0.1, n_pix_separation=8, n_pix_edge=4, nTheta_user_defined=16)
    N = wfs.nRes
    wfs.focal_plane_camera.resolution = N
    basis = dm.modes.reshape(32, 32, dm.nValidAct)[:, :, :20] * tel.pupil[:, :, None]
    gsc = GainSensingCamera(wfs.mask, basis)
    tel.resetOPD()
    ngs * tel * wfs
    wfs * wfs.focal_plane_camera
    f0 = wfs.focal_plane_camera.frame.copy()
    wfs.focal_plane_camera * gsc                          # calibration
    assert gsc.calibration_ready and np.all(gsc.og == 1)
    dm.coefs = np.random.RandomState(5).normal(0, 60e-9, dm.nValidAct)
    ngs * tel * dm * wfs
    wfs * wfs.focal_plane_camera
    f1 = wfs.focal_plane_camera.frame.copy()
    wfs.focal_plane_camera * gsc                          # optical gains
    o = GSCOracle(wfs.mask, basis)
    o.calibration(f0)
    og_ref = o.optical_gains(f1)
    e_bp = rel_l2(gsc.basis_product, o.bp.real)
    e_ir = rel_l2(gsc.IR_calib, o.ir_calib)
    e_og = np.max(np.abs(gsc.og - og_ref)) / np.max(np.abs(og_ref))
    print('basis product %.2e  IR_calib %.2e  og %.2e' % (e_bp, e_ir, e_og), 'og', np.round(gsc.og[:6], 4))
    assert e_bp < 1e-5 and e_ir < 1e-5 and e_og < 1e-4
    # recalibration resets the gains; the same frame gives og = 1
    gsc.reset_calibration()
    assert not gsc.calibration_ready
    wfs.focal_plane_camera * gsc
    wfs.focal_plane_camera * gsc
    assert np.max(np.abs(gsc.og - 1)) < 1e-5
