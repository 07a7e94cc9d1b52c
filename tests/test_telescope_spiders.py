"""Telescope.apply_spiders (Telescope.py:425-459) on the host pupil: each spider removes a strip of the
given width through the (offset) centre on one side only; checked against the strip geometry written
out directly on the pixel grid."""
import numpy as np


def _tel(n=64):
    from oopao_amd import Telescope
    return Telescope(n, 8.0, 1e-3, centralObstruction=0.2, device='cpu')


def test_single_spider_is_a_half_strip():
    tel = _tel()
    base = tel.pupil.copy()
    tel.apply_spiders([0], 0.5)
    x = np.linspace(-4, 4, 64)
    X, Y = np.meshgrid(x, x)
    strip = (np.abs(Y) <= 0.25) & (np.arange(64)[None, :] >= 32)     # angle 0 -> towards +x
    assert np.array_equal(tel.pupil, base & ~strip)
    assert tel.pixelArea == int((base & ~strip).sum())


def test_four_spiders_with_offset_and_reset():
    tel = _tel()
    base = tel.pupil.copy()
    tel.apply_spiders([0, 90, 180, 270], 0.3, offset_X=[0.1] * 4, offset_Y=[0.0] * 4)
    x = np.linspace(-4, 4, 64)
    X, Y = np.meshgrid(x + 0.1, x)
    i = np.arange(64)
    cut = ((np.abs(Y) <= 0.15) & (i[None, :] >= 32)) | ((np.abs(X) <= 0.15) & (i[:, None] >= 32)) | \
          ((np.abs(Y) <= 0.15) & (i[None, :] < 32)) | ((np.abs(X) <= 0.15) & (i[:, None] < 32))
    assert np.array_equal(tel.pupil, base & ~cut)
    tel.apply_spiders([0], 0.0)                                  # thickness <= 0: default pupil back
    assert np.array_equal(tel.pupil, base)
