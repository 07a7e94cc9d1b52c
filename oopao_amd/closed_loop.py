"""Integrator closed loop with all state on the GPU.

Same sequence as tutorials/AO_closed_loop_Pyramid_WFS.py:601-638 (and
closed_loop/run_cl.py:176-217): per iteration
    atm.update();  ngs**atm*tel*dm*wfs
    frame_delay == 1: s = wfs.signal
    dm.coefs = dm.coefs - gain * reconstructor @ s
    frame_delay == 2: s = wfs.signal
with the OPD sum, PWFS frame, slopes, reconstructor GEMV and integrator update
kept on the device (no host round trip per iteration). The reconstructor is the
matrix R [nAct, nSignal] the tutorial forms (reconstructor = M2C_CL @ calib_CL.M),
or the pair (M2C [nAct, nModes], M [nModes, nSignal]) of its factors: then
R @ s = M2C @ (M @ s), two GEMVs moving (nAct + nSignal) nModes values per
iteration instead of nAct nSignal (61 MB instead of 210 MB in float32 at 1000
modes of the 80 x 80 sensor), the same update up to float rounding. With
``reconstructor_precision='bf16'`` the reconstructor is held in BF16 (half the
bytes per iteration, FP32 accumulation), with ``'bf16x3'`` as split BF16 (hi + lo,
the reconstructor to 2^-16, the float32 loop's accuracy). Records the
turbulence and residual wavefront error in the pupil (nm RMS) per iteration,
and, with ``strehl_every`` > 0, the Strehl ratio of the source's image behind
the DM (ngs*tel*dm*cam with strehlMeter against the diffraction-limited image,
tutorials/AO_closed_loop_Pyramid_WFS.py:609-614), computed on the GPU.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._device import ptr, stream_ptr


def run_closed_loop(tel, ngs, atm, dm, wfs, reconstructor, gain=0.4, n_iter=100, frame_delay=2,
                    record_slopes=False, strehl_every=0, psf_sampling=2, reconstructor_precision='fp32'):
    dev = wfs.device

    def as_dev(a):
        a = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32))
        return a.to(dev, torch.float32).contiguous()
    factored = isinstance(reconstructor, (tuple, list))
    if factored:
        if len(reconstructor) != 2:
            raise ValueError('a factored reconstructor is the pair (M2C, M), got %d factors' % len(reconstructor))
        F1, F2 = as_dev(reconstructor[0]), as_dev(reconstructor[1])
        if F1.dim() != 2 or F2.dim() != 2 or F1.shape[1] != F2.shape[0]:
            raise ValueError('factors %s and %s do not chain' % (tuple(F1.shape), tuple(F2.shape)))
        nAct, nSig, nModes = F1.shape[0], F2.shape[1], F1.shape[1]
        mode_vec = torch.empty(nModes, dtype=torch.float32, device=dev)
    else:
        R = as_dev(reconstructor)
        nAct, nSig = R.shape
    if nAct != dm.nValidAct or nSig != wfs.nSignal:
        raise ValueError('reconstructor shape %s does not match (nAct=%d, nSignal=%d)'
                         % ((nAct, nSig), dm.nValidAct, wfs.nSignal))
    if frame_delay not in (1, 2):
        raise ValueError('frame_delay must be 1 or 2, got %r' % (frame_delay,))
    if reconstructor_precision not in ('fp32', 'bf16', 'bf16x3'):
        raise ValueError("reconstructor_precision must be 'fp32', 'bf16' or 'bf16x3', got %r"
                         % (reconstructor_precision,))
    recon_bf16 = None
    if reconstructor_precision != 'fp32':
        from .contract import Contraction, PackedBF16
        split = reconstructor_precision == 'bf16x3'
        if factored:
            recon_bf16 = [Contraction(PackedBF16(F1, split=split)), Contraction(PackedBF16(F2, split=split))]
        else:
            recon_bf16 = Contraction(PackedBF16(R, split=split))
    pupil = tel._pupil_dev
    pupil_f = (pupil > 0).float().contiguous()
    npix = int((pupil > 0).sum())
    nb = ctypes.c_size_t(0)
    _lib.call('oopao_loop_residual_workspace_bytes', ctypes.byref(nb))
    resid_ws = torch.empty(nb.value, dtype=torch.uint8, device=dev)
    opd = torch.empty((tel.resolution, tel.resolution), dtype=torch.float32, device=dev)
    coefs = torch.as_tensor(np.asarray(dm.coefs, np.float32), device=dev).clone()
    new_coefs = torch.empty_like(coefs)
    wfe_atm = torch.zeros(n_iter, dtype=torch.float64, device=dev)
    wfe_res = torch.zeros(n_iter, dtype=torch.float64, device=dev)
    slopes = torch.zeros((n_iter, nSig), dtype=torch.float32, device=dev) if record_slopes else None
    tel.isPaired = False
    ngs ** tel
    amp = wfs._amp()
    scale = 2 * np.pi / ngs.wavelength
    eng, se = wfs.engine, wfs.slopes_engine
    buf = eng.buffers(1)
    bands = wfs.slope_bands()         # the loop reads only the slopes: quadrant columns of the frame
    if strehl_every:
        from .science import psf_dev, strehl_dev
        amp_psf = (torch.sqrt(ngs._intensity) * pupil).float().contiguous()
        psf_scale = 2 * np.pi / ngs.wavelength
        zero = torch.zeros((tel.resolution, tel.resolution), dtype=torch.float32, device=dev)
        psf_ref = psf_dev(amp_psf, zero, psf_scale, tel.resolution, psf_sampling, None, dev)
        n_sr = (n_iter + strehl_every - 1) // strehl_every
        sr = torch.zeros(n_sr, dtype=torch.float64, device=dev)
    frame = torch.empty((1, eng.cam, eng.cam), dtype=torch.float32, device=dev)
    s2d = torch.empty((1, 2 * se.w, se.w), dtype=torch.float32, device=dev)
    # slope vectors of this and of the previous iteration (swapped, no copy): the 2-frame delay applies
    # the previous one
    s_cur = torch.zeros((1, se.nsig), dtype=torch.float32, device=dev)
    s_prev = torch.zeros((1, se.nsig), dtype=torch.float32, device=dev)
    for i in range(n_iter):
        atm.update()
        turb = atm._OPD_dev
        dm_opd = dm.opd_from_coefs(coefs)
        # opd = (turb + dm) * pupil and both RMS errors, written into the traces on the device
        _lib.call('oopao_loop_residual', ptr(turb), ptr(dm_opd), ptr(pupil_f), ptr(opd), opd.numel(),
                  ctypes.c_double(npix), ptr(resid_ws), ptr(wfe_atm[i:]), ptr(wfe_res[i:]), stream_ptr(dev))
        if strehl_every and i % strehl_every == 0:
            psf = psf_dev(amp_psf, opd, psf_scale, tel.resolution, psf_sampling, None, dev)
            sr[i // strehl_every] = strehl_dev(psf, psf_ref, tel.resolution, psf_sampling, as_tensor=True)
        eng.rows(opd[None], scale, amp, buf, bands)
        eng.cols(buf, frame, bands)
        se.slopes(frame, s2d, s_cur)
        if record_slopes:
            slopes[i] = s_cur[0]
        sig = s_cur[0] if frame_delay == 1 else s_prev[0]
        # coefs <- coefs - gain * R @ sig   (R [nAct, nSig] row-major GEMV, float32 or BF16 matrix; factored:
        # the modal vector M @ sig first)
        if factored and recon_bf16 is not None:
            recon_bf16[1].gemv(sig, out=mode_vec, precision=reconstructor_precision)
            recon_bf16[0].gemv(mode_vec, y0=coefs, alpha=-gain, beta=1.0, out=new_coefs,
                               precision=reconstructor_precision)
        elif factored:
            _lib.call('oopao_gemv', ptr(F2), ptr(sig), ptr(mode_vec), None, nModes, nSig, 1.0, 0.0, stream_ptr(dev))
            _lib.call('oopao_gemv', ptr(F1), ptr(mode_vec), ptr(new_coefs), ptr(coefs), nAct, nModes, float(-gain),
                      1.0, stream_ptr(dev))
        elif recon_bf16 is not None:
            recon_bf16.gemv(sig, y0=coefs, alpha=-gain, beta=1.0, out=new_coefs, precision=reconstructor_precision)
        else:
            _lib.call('oopao_gemv', ptr(R), ptr(sig), ptr(new_coefs), ptr(coefs), nAct, nSig, float(-gain), 1.0,
                      stream_ptr(dev))
        coefs, new_coefs = new_coefs, coefs
        if frame_delay == 2:
            s_cur, s_prev = s_prev, s_cur
    torch.cuda.synchronize(dev)
    dm.coefs = coefs.double().cpu().numpy()
    out = dict(wfe_atmosphere=wfe_atm.cpu().numpy(), wfe_residual=wfe_res.cpu().numpy(),
               coefs=coefs.double().cpu().numpy())
    if record_slopes:
        out['slopes'] = slopes.double().cpu().numpy()
    if strehl_every:
        out['strehl'] = sr.cpu().numpy()
        out['strehl_every'] = strehl_every
    return out
