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
    return measure_opd(tel, wfs, opd)


def measure_opd(tel, wfs, opd):
    """Slopes [nSignal, B] (device float32) of the OPD maps opd [B, n, n] (device, metres) through tel*wfs."""
    pupil = tel._pupil_dev
    x = (opd * pupil).contiguous()
    # the calibration keeps only the slopes: with an identity camera readout the slopesMaps signal reads
    # the quadrant columns only (as in Pyramid.wfs_measure), and the camera's whole frame is computed if read
    bands = None
    if wfs.isInitialized and wfs.isCalibrated and wfs.cam.identity_readout():
        bands = wfs.slope_bands()
    frames = wfs.frames_dev(x, bands=bands)
    if wfs.binning != 1:
        b = wfs.binning
        c = frames.shape[-1]
        frames = frames.reshape(-1, c // b, b, c // b, b).sum(dim=(2, 4))
    wfs.cam.integrate(frames)
    if bands is not None:
        wfs.cam._defer_full(lambda: wfs.frames_dev(x))
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


def InteractionMatrixFromPhaseScreen(ngs, atm, tel, wfs, phasScreens, stroke, phaseOffset=0, nMeasurements=50,
                                     noise='off', invert=True, print_time=True):
    """Push-pull interaction matrix of OPD maps (calibration/InteractionMatrix.py:172-259).

    phasScreens is [n, n, nModes] (or one [n, n] map), numpy or a device tensor, in the reference's
    layout; column k is 0.5 (s(+stroke screen_k) - s(-stroke screen_k)) / stroke. phaseOffset (radians,
    [n, n]) is added to the phase of both the push and the pull maps as the reference's
    tel.src.phase += phaseBuffer (InteractionMatrix.py:225-232). Each batch of nMeasurements maps goes
    through the WFS as one batched device call; the columns accumulate in a device float64 matrix.
    """
    if noise == 'off':
        wfs.cam.photonNoise = 0
        wfs.cam.readoutNoise = 0
    else:
        print('Warning: Keeping the noise configuration for the WFS')
    tel.isPaired = False
    ngs * tel
    dev = wfs.device
    scr = phasScreens if isinstance(phasScreens, torch.Tensor) else torch.from_numpy(np.asarray(phasScreens))
    if scr.dim() == 2:
        scr = scr[:, :, None]
    if scr.dim() != 3 or scr.shape[0] != tel.resolution or scr.shape[1] != tel.resolution:
        raise OopaoError('phasScreens must be [%d, %d, nModes], got %s'
                         % (tel.resolution, tel.resolution, tuple(scr.shape)))
    scr = scr.to(dev, torch.float32).permute(2, 0, 1)   # [nModes, n, n] view
    nModes = scr.shape[0]
    nMeasurements = max(1, min(int(nMeasurements), nModes))
    offset = None
    if np.ndim(phaseOffset) == 2:
        off = phaseOffset if isinstance(phaseOffset, torch.Tensor) else torch.from_numpy(np.asarray(phaseOffset))
        offset = (off.to(dev, torch.float64) * (ngs.wavelength / (2 * np.pi))).to(torch.float32)
    elif np.ndim(phaseOffset) == 0 and phaseOffset != 0:
        offset = torch.full((tel.resolution, tel.resolution), float(phaseOffset) * ngs.wavelength / (2 * np.pi),
                            dtype=torch.float32, device=dev)
    intMat = torch.zeros((wfs.nSignal, nModes), dtype=torch.float64, device=dev)
    for i0 in range(0, nModes, nMeasurements):
        a = time.time()
        i1 = min(nModes, i0 + nMeasurements)
        push = scr[i0:i1] * stroke
        pull = -push
        if offset is not None:
            push = push + offset
            pull = pull + offset
        sp = measure_opd(tel, wfs, push.contiguous())
        sm = measure_opd(tel, wfs, pull.contiguous())
        intMat[:, i0:i1] = 0.5 * (sp.double() - sm.double()) / stroke
        if print_time:
            torch.cuda.synchronize(dev)
            print(str(i1) + '/' + str(nModes))
            print('Time elapsed: ' + str(time.time() - a) + ' s')
    ngs * tel
    out = CalibrationVault(intMat, invert=invert, device=dev)
    out.D_dev = intMat
    return out
