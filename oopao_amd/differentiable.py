"""Differentiable modulated PWFS: torch.autograd.Function over the HIP forward and its adjoint.

The forward is the engine's frame kernel (oopao_pwfs_frames on the separable path,
oopao_pwfs_generic_frames for any other N, mask or binning); the backward is the matching hand-written
adjoint (oopao_pwfs_adjoint / oopao_pwfs_generic_adjoint): w_i = M^H (g . E_i) with the conjugated mask,
and dL/dphi = sum_i 2 Im(w_i conj(u_i)), recomputing every modulation point instead of storing it.
The slopesMaps post-processing (Pyramid.py:835-854, grabQuadrant 946-961) is written as torch
operations so that autograd carries the gradient from the slopes to the detector frame.

There is no counterpart in the reference (OOPAO has no gradients); the forward it differentiates is
Pyramid.pyramid_transform / wfs_measure / signalProcessing (Pyramid.py:612-646, 712-757, 835-854).
"""
import numpy as np
import torch

from .tools import OopaoError


class PWFSFrames(torch.autograd.Function):
    """frames [B, cam, cam] = PWFS(opd [B, n, n]); phase = opd * scale. With ``bands`` (detector-column
    intervals, separable path) only those columns are computed, the others are zero, and the backward
    treats the cotangent as zero there: for a downstream that reads only those columns (the slopesMaps
    signal reads the quadrant columns) the values and gradients are those of the whole frame."""

    @staticmethod
    def forward(ctx, opd, engine, amp, scale, bands=None):
        opd = opd.detach().to(torch.float32).contiguous()
        frames = engine.frames(opd, scale, amp, bands=bands)
        ctx.save_for_backward(opd)
        ctx.engine, ctx.amp, ctx.scale, ctx.bands = engine, amp, scale, bands
        return frames

    @staticmethod
    def backward(ctx, gframes):
        opd, = ctx.saved_tensors
        grad = ctx.engine.adjoint(opd, ctx.scale, ctx.amp, gframes.detach().to(torch.float32).contiguous(),
                                  bands=ctx.bands)
        return grad, None, None, None, None


def pwfs_frames(engine, opd, amp, scale, bands=None):
    """Differentiable detector frames for OPD/phase maps opd [B, n, n] (``bands``: see PWFSFrames)."""
    return PWFSFrames.apply(opd, engine, amp, scale, bands)


def slopes_maps(frames, validI4Q, quad, w, ref2d=0.0, units=1.0):
    """slopesMaps of frames [B, cam, cam] in torch ops (Pyramid.py:838-854): returns
    (signal_2D [B, 2w, w], signal [B, nSig])."""
    v = validI4Q.to(frames.dtype)
    I = [frames[:, quad[k]:quad[k] + w, quad[4 + k]:quad[4 + k] + w] * v for k in range(4)]
    I1, I2, I3, I4 = I
    I4Q = I1 + I2 + I3 + I4
    norma = (I4Q.double().sum(dim=(1, 2)) / v.double().sum()).to(frames.dtype)[:, None, None]
    Sx = I1 - I2 + I4 - I3
    Sy = I1 - I4 + I2 - I3
    s2d = (torch.cat((Sx, Sy), dim=1) / norma - ref2d) * units
    vs = torch.cat((validI4Q, validI4Q), dim=0).bool()
    return s2d, s2d[:, vs]


def sensor_frames(wfs, phase, bands=None):
    """Differentiable detector frames of a drop-in Pyramid or Bi-O Edge for phase maps [B, n, n] in
    radians (before the detector's extra binning): the Pyramid's frame, or the Bi-O Edge's four mask
    images tiled [[A, B], [C, D]] (BioEdge.py:510-544). ``bands`` (Pyramid on the separable path):
    compute only those detector columns (see PWFSFrames)."""
    amp = wfs._amp()
    engines = getattr(wfs, 'engines', None)
    if engines is None:
        return pwfs_frames(wfs.engine, phase, amp, 1.0, bands)
    t = [pwfs_frames(e, phase, amp, 1.0) for e in engines]
    return torch.cat((torch.cat((t[0], t[1]), dim=2), torch.cat((t[2], t[3]), dim=2)), dim=1)


def pwfs_signal(wfs, phase):
    """Differentiable slopes [B, nSignal] of a drop-in Pyramid or Bi-O Edge (slopesMaps post-processing)
    for phase maps [B, n, n] in radians at the WFS wavelength; gradients flow to ``phase``. Any sensor
    geometry of the separable or the general path (any N, mask, integer binning) is covered."""
    if wfs.postProcessing != 'slopesMaps':
        raise OopaoError('pwfs_signal supports the slopesMaps post-processing')
    if wfs.engine.fractional:
        raise OopaoError('the PWFS adjoint does not cover non-integer detector sampling')
    phase = phase.reshape(-1, wfs.engine.n, wfs.engine.n)
    # the slopes read only the quadrant columns: the forward computes those columns and the adjoint
    # transforms only them (a Pyramid on the separable path; the Bi-O Edge tiles four frames)
    bands = wfs.slope_bands() if getattr(wfs, 'engines', None) is None else None
    frames = sensor_frames(wfs, phase, bands)
    if wfs.binning != 1:                       # the detector's extra binning (Pyramid.py:800-808)
        b, c = wfs.binning, frames.shape[-1]
        frames = frames.reshape(-1, c // b, b, c // b, b).sum(dim=(2, 4))
    quad, w = wfs._quadrant_origins()
    valid = torch.from_numpy(np.asarray(wfs.validI4Q, bool)).to(frames.device)
    ref = wfs._referenceSignal_2D if wfs._referenceSignal_2D is not None else 0.0
    return slopes_maps(frames, valid, quad, w, ref, float(wfs.slopesUnits))[1]


def _vjp_batch(wfs, phase0, G):
    """J(phase0)^T g_k for the slope cotangents G [K, nSignal]: one forward + one adjoint pass over K
    copies of the working point; returns [K, n, n] (radians^-1 per slope unit)."""
    n = wfs.engine.n
    ph = phase0.reshape(1, n, n).to(torch.float32).expand(G.shape[0], n, n).contiguous().requires_grad_(True)
    s = pwfs_signal(wfs, ph)
    (s * G.to(s.dtype)).sum().backward()
    return ph.grad


def slopes_jacobian(wfs, phase0, basis, batch=64):
    """Jacobian J [nSignal, nModes] = d signal / d a of the slopes w.r.t. the coefficients a of
    ``basis`` [nModes, n, n] (phase in radians per unit coefficient) at the working point
    ``phase0`` [n, n] (radians): the interaction matrix of the linearised sensor at phase0, from
    nSignal vector-Jacobian products (unit slope cotangents) through the adjoint, batched by
    ``batch``. At phase0 = 0 it is the push-pull interaction matrix in the small-stroke limit."""
    dev = wfs.device
    nsig = int(wfs.nSignal)
    B = basis.reshape(basis.shape[0], -1).to(dev, torch.float32)
    J = torch.empty((nsig, B.shape[0]), dtype=torch.float32, device=dev)
    for i0 in range(0, nsig, batch):
        i1 = min(nsig, i0 + batch)
        G = torch.zeros((i1 - i0, nsig), dtype=torch.float32, device=dev)
        G[torch.arange(i1 - i0, device=dev), torch.arange(i0, i1, device=dev)] = 1.0
        g = _vjp_batch(wfs, phase0, G).reshape(i1 - i0, -1)
        J[i0:i1] = g @ B.t()
    return J


def optical_gains_adjoint(wfs, phase0, basis, D_calib, batch=64):
    """Modal optical gains at the working point phase0 [n, n] (radians) with respect to the
    calibration interaction matrix D_calib [nSignal, nModes] (slopes per unit coefficient of
    ``basis`` [nModes, n, n], phase in radians):
        og_m = <D_calib[:, m], J(phase0) b_m> / <D_calib[:, m], D_calib[:, m]>,
    the projection of each mode's sensitivity at phase0 on its calibrated response, with
    J(phase0)^T D_calib[:, m] from one adjoint pass per mode (batched)."""
    dev = wfs.device
    Bm = basis.reshape(basis.shape[0], -1).to(dev, torch.float32)
    D = D_calib.to(dev, torch.float32)
    M = Bm.shape[0]
    og = torch.empty(M, dtype=torch.float64, device=dev)
    for m0 in range(0, M, batch):
        m1 = min(M, m0 + batch)
        G = D[:, m0:m1].t().contiguous()
        g = _vjp_batch(wfs, phase0, G).reshape(m1 - m0, -1)
        num = (g.double() * Bm[m0:m1].double()).sum(dim=1)
        og[m0:m1] = num / (G.double() ** 2).sum(dim=1)
    return og
