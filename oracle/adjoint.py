"""numpy float64 adjoint of the PWFS forward model (CPU oracle, test infrastructure only).

OOPAO has no gradients; this differentiates the oracle's restatement of Pyramid.pyramid_transform
(Pyramid.py:612-646), the modulation sum and sum-binning (712-757, tools/tools.py:464-483) and the
slopesMaps post-processing (835-854) directly from their definitions, with full N x N FFTs:

  E_i = ifft2(fft2(embed(u_i) Phi) mask),  u_i = amp exp(i (phi + TT_i)),
  frame = bin(sum_i |E_i|^2),
  dL/dphi = sum_i 2 Im(w_i conj(u_i)),  w_i = crop(conj(Phi) ifft2(conj(mask) fft2(g E_i))),
  g = dL/dframe repeated over each b x b bin.

Pins (tests/test_oracle_adjoint.py): central finite differences of the oracle forward in float64 along
random directions, for the frame VJP and for the slopes VJP.
"""
import numpy as np
import scipy.fft

from .pwfs import grab_quadrant


def frame_vjp(o, phase, gframe):
    """dL/dphase [n, n] for dL/dframe = gframe [cam, cam] at src.phase = phase (oracle ``o``)."""
    g = o.geom
    n, N, c = g.n, g.N, g.center
    b = N // g.cam
    gup = np.repeat(np.repeat(np.asarray(gframe, np.float64), b, axis=0), b, axis=1)
    amp = np.sqrt(o.flux.astype(np.float64) / o.nTheta)
    ph = np.asarray(phase, np.float64)
    tts = [0.0] if o.tt is None else [o.tt[i].astype(np.float64) for i in range(o.nTheta)]
    mask = o.mask.astype(np.complex128)
    Phi = o.phasor
    grad = np.zeros((n, n))
    sl = slice(c - n // 2, c + n // 2)
    for tt in tts:
        u = amp * np.exp(1j * (ph + tt))
        S = np.zeros((N, N), complex)
        S[sl, sl] = u
        E = np.fft.ifft2(scipy.fft.fft2(S * Phi) * mask)
        Y = gup * E
        W = np.conj(Phi) * np.fft.ifft2(np.conj(mask) * np.fft.fft2(Y))
        w = W[sl, sl]
        grad += 2 * np.imag(w * np.conj(u))
    return grad


def slopes_vjp(o, frame, dsig):
    """dL/dframe [cam, cam] for dL/dsignal = dsig [nSig] of slopes_maps(frame) (Pyramid.py:838-854)."""
    g = o.geom
    v = np.asarray(o.validI4Q, bool)
    w = v.shape[0]
    I = [grab_quadrant(frame, k, g) * v for k in (1, 2, 3, 4)]
    I1, I2, I3, I4 = I
    nvalid = v.sum()
    norma = np.sum((I1 + I2 + I3 + I4)[v]) / nvalid
    Sx = I1 - I2 + I4 - I3
    Sy = I1 - I4 + I2 - I3
    D = np.zeros((2 * w, w))
    D[np.concatenate((v, v))] = dsig
    dSx, dSy = D[:w], D[w:]
    dnorma = -np.sum(dSx * Sx + dSy * Sy) / norma ** 2
    common = dnorma / nvalid
    dI = [(dSx + dSy) / norma + common, (-dSx + dSy) / norma + common,
          (-dSx - dSy) / norma + common, (dSx - dSy) / norma + common]
    out = np.zeros_like(frame, dtype=np.float64)
    for k, d in zip((1, 2, 3, 4), dI):
        view = grab_quadrant(out, k, g)
        view += d * v
    return out
