"""numpy restatement of the science path (CPU oracle, test infrastructure only).

Telescope.PropagateField (Telescope.py:374-423) for an on-axis source without coronagraph, and
strehlMeter (tools/tools.py:102-157).

Pins (tests/test_oracle_science.py): Parseval (the whole padded image holds the pupil energy), the
diffraction-limited image's point symmetry about the centre of its four central pixels, a unit
Strehl for the diffraction-limited image, and the Marechal approximation exp(-sigma^2) for small
random phases.
"""
import math

import numpy as np


def propagate_field(amplitude, phase, zeroPaddingFactor=2, img_resolution=None):
    """|EMF|^2 on the centred img x img window (oversampling binned), float64."""
    n = amplitude.shape[0]
    over = 1
    if zeroPaddingFactor * over < 2:
        over = int(math.ceil(2.0 / zeroPaddingFactor))
    if img_resolution is None:
        img_resolution = zeroPaddingFactor * n
    img = int(math.ceil(img_resolution * over))
    N = int(np.fix(zeroPaddingFactor * over * n))
    pad = int(math.ceil((N - n) / 2))
    S = np.pad(amplitude * np.exp(1j * phase), ((pad, pad), (pad, pad)))
    N = S.shape[0]
    xx, yy = np.meshgrid(np.arange(N), np.arange(N))
    ph = np.exp(-1j * np.pi * (N + 1) / N * (xx + yy) * (1 - img % 2))
    EMF = np.fft.fft2(S * ph) / N
    if N % 2 == img % 2:
        shift = 0
    else:
        shift = 1 if N % 2 == 0 else -1
    i0 = int(np.ceil(N / 2) - img // 2 + (1 - N % 2) - 1)
    i1 = int(np.ceil(N / 2) + img // 2 + shift)
    psf = np.abs(EMF[i0:i1, i0:i1]) ** 2
    if over != 1:
        r = psf.shape[0] // over
        psf = psf.reshape(r, over, r, over).sum(axis=(1, 3))
    return psf


def strehl(psf, psf_ref):
    """Sum of the normalised OTF over that of the reference image, in % (tools/tools.py:114-157)."""
    s = min(psf.shape[0], psf_ref.shape[0])

    def crop(a):
        c = a.shape[0] // 2
        return a[c - s // 2:c + s // 2, c - s // 2:c + s // 2]
    otf = np.abs(np.fft.fft2(crop(psf)))
    otfa = np.abs(np.fft.fft2(crop(psf_ref)))
    return float((otf / otf.max()).sum() / (otfa / otfa.max()).sum() * 100)
