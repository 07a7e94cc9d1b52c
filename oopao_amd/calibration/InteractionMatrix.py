"""Push(-pull) interaction matrix (calibration/InteractionMatrix.py:13-169), batched on the GPU.

For each batch of modes the DM OPD cube (modes @ M2C[:, batch] * stroke), the
modulated PWFS frames of every mode and the slopes are computed on the device
(oopao_amd/csrc/dm.hip, pwfs.hip); the columns accumulate in a device float64
matrix handed to CalibrationVault. Column convention as the reference:
single pass -> D[:, k] = s_push / stroke; push-pull -> 0.5 (s_push - s_pull) / stroke.
"""
import time

import numpy as np
import torch

from .CalibrationVault import CalibrationVault
from ..tools import OopaoError


def measure_modes(ngs, tel, dm, wfs, coefs):
    """Slopes [nSignal, B] (device float32) of the DM commands coefs [nAct, B] (device) through ngs**tel*dm*wfs."""
    opd = dm.opd_from_coefs(coefs)                     # [B, n, n]
    ngs ** tel
    pupil = tel._pupil_dev
    frames = wfs.frames_dev((opd * pupil).contiguous())
    if wfs.binning != 1:
        b = wfs.binning
        c = frames.shape[-1]
        frames = frames.reshape(-1, c // b, b, c // b, b).sum(dim=(2, 4))
    wfs.cam.integrate(frames)
    s2d, s = wfs._signal_processing_dev()
    return s.t()


def InteractionMatrix(ngs, tel, dm, wfs, M2C, stroke, atm=None, phaseOffset=0, nMeasurements=50, noise='off',
                      invert=True, nTrunc=0, print_time=False, display=False, single_pass=True):
    """Interaction matrix of `M2C` modes (columns) through the WFS; returns a CalibrationVault."""
    if noise == 'off':
        wfs.cam.photonNoise = 0
        wfs.cam.readoutNoise = 0
        wfs.cam.backgroundNoise = 0
    else:
        print('Warning: Keeping the noise configuration for the WFS')
    if np.ndim(phaseOffset) == 2 and np.any(phaseOffset != 0):
        raise OopaoError('a phase offset is not supported by oopao_amd (it does not affect the push measurement '
                         'of the reference either, InteractionMatrix.py:120-123)')
    tel.isPaired = False
    ngs ** tel
    dev = wfs.device
    M2C_t = M2C if isinstance(M2C, torch.Tensor) else torch.from_numpy(np.asarray(M2C, dtype=np.float64))
    if M2C_t.dim() == 1:
        M2C_t = M2C_t[:, None]
    M2C_t = M2C_t.to(dev, torch.float32)
    nModes = M2C_t.shape[1]
    nMeasurements = max(1, min(int(nMeasurements), nModes))
    intMat = torch.zeros((wfs.nSignal, nModes), dtype=torch.float64, device=dev)
    t0 = time.time()
    for i0 in range(0, nModes, nMeasurements):
        i1 = min(nModes, i0 + nMeasurements)
        cmd = M2C_t[:, i0:i1] * stroke
        sp = measure_modes(ngs, tel, dm, wfs, cmd)
        if single_pass:
            col = sp.double() / stroke
        else:
            sm = measure_modes(ngs, tel, dm, wfs, -cmd)
            col = 0.5 * (sp.double() - sm.double()) / stroke
        intMat[:, i0:i1] = col
        if print_time:
            print(str(i1) + '/' + str(nModes))
    if print_time:
        torch.cuda.synchronize()
        print('Time elapsed: ' + str(time.time() - t0) + ' s')
    dm.coefs = 0
    ngs ** tel
    out = CalibrationVault(intMat, invert=invert, nTrunc=nTrunc, device=dev)
    out.D_dev = intMat
    return out
