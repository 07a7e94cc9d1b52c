"""CPU: pins of the science-path oracle (oracle/science.py; Telescope.py:374-423, tools/tools.py:102-157)."""
import numpy as np

from oracle import science as osc
from oracle.pwfs import telescope_pupil
from tests.helpers import random_phase


def test_psf_parseval_symmetry_and_unit_strehl():
    n = 32
    pup = telescope_pupil(n).astype(float)
    amp = pup * 3.0
    full = osc.propagate_field(amp, np.zeros((n, n)), 2, 64)
    assert abs(full.sum() / (amp ** 2).sum() - 1) < 1e-12
    # real field: point-symmetric image about the centre of the four central pixels
    assert np.max(np.abs(full - full[::-1, ::-1])) < 1e-12 * full.max()
    peak = np.unravel_index(np.argmax(full), full.shape)
    assert peak[0] in (31, 32) and peak[1] in (31, 32)
    crop = osc.propagate_field(amp, np.zeros((n, n)), 2, 40)
    assert np.allclose(crop, full[12:52, 12:52])
    assert abs(osc.strehl(full, full) - 100) < 1e-12


def test_strehl_marechal_small_phase():
    n = 64
    pup = telescope_pupil(n).astype(float)
    ref = osc.propagate_field(pup, np.zeros((n, n)), 2)
    for rms in (0.2, 0.4):
        ph = random_phase(n, 9, rms, pup)
        sel = ph[pup > 0]
        var = np.var(sel)
        sr = osc.strehl(osc.propagate_field(pup, ph, 2), ref) / 100
        assert abs(sr - np.exp(-var)) < 0.05 * np.exp(-var), (sr, np.exp(-var))
