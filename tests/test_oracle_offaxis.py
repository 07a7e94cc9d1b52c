"""CPU: pins of the off-axis / laser-guide-star pieces of the atmosphere oracle (oracle/atmosphere.py)."""
import os

import numpy as np

from oracle.atmosphere import source_layer, warp_bilinear

GOLD = os.path.join(os.path.dirname(__file__), 'golden', 'bilinear_skimage.npz')


def test_warp_bilinear_matches_skimage():
    g = np.load(GOLD)
    for k in range(int(g['n_cases'])):
        img, par, out = g['img_%d' % k], g['par_%d' % k], g['out_%d' % k]
        if int(g['kind_%d' % k]) == 0:
            got = warp_bilinear(img, 1.0, -par[1], -par[0], img.shape)
        else:
            ratio, extra = par
            c = img.shape[0] / 2
            b = -(c + extra) / ratio + c
            got = warp_bilinear(img, 1.0 / ratio, b, b, img.shape)
        assert np.max(np.abs(got - out)) < 1e-12, k


def _ramp(res, ax=0.37, ay=-0.21, c0=1.5):
    r, c = np.meshgrid(np.arange(res, dtype=float), np.arange(res, dtype=float), indexing='ij')
    return ax * c + ay * r + c0, (ax, ay, c0)


def test_source_layer_closed_forms():
    n, res, D = 16, 40, 40 * 8.0 / 16
    im, (ax, ay, c0) = _ramp(res)
    rad2arcsec = 180 / np.pi * 3600
    # on-axis source at infinity: the centred crop
    assert np.array_equal(source_layer(im, n, D, 5000.0, 0.0, 0.0, np.inf), im[res // 2 - n // 2:res // 2 + n // 2,
                                                                               res // 2 - n // 2:res // 2 + n // 2])
    # off-axis NGS: the ramp evaluated at the footprint's continuous position (bilinear is exact on a ramp)
    h, z, az = 5000.0, 30.0, 35.0
    rho = h * np.tan(z / rad2arcsec) * res / D
    x_z, y_z = rho * np.cos(np.deg2rad(az)), rho * np.sin(np.deg2rad(az))
    r, c = np.meshgrid(np.arange(n, dtype=float), np.arange(n, dtype=float), indexing='ij')
    X = c + res // 2 - n // 2 + x_z          # continuous column position on the layer
    Y = r + res // 2 - n // 2 + y_z
    assert np.allclose(source_layer(im, n, D, h, z, az, np.inf), ax * X + ay * Y + c0, rtol=0, atol=1e-12)
    # LGS on axis: the centred crop magnified about the footprint centre by H / (H - h)
    H = 90e3
    m = (H - h) / H
    ratio = 1 / m
    extra = (ratio - 1) / 2
    u = (c - n / 2 - extra) / ratio + n / 2
    v = (r - n / 2 - extra) / ratio + n / 2
    ref = ax * (u + res // 2 - n // 2) + ay * (v + res // 2 - n // 2) + c0
    got = source_layer(im, n, D, h, 0.0, 0.0, H)
    inner = (slice(1, n - 1), slice(1, n - 1))     # bilinear is exact away from the crop edge
    assert np.allclose(got[inner], ref[inner], rtol=0, atol=1e-12)
    # the pixel-centre alignment: the magnification keeps the centre (n - 1) / 2 fixed
    assert np.isclose((((n - 1) / 2) - n / 2 - extra) / ratio + n / 2, (n - 1) / 2)


def test_asm_gaussian_beam_parseval_identity():
    from oracle.atmosphere import asm
    N, pitch, wvl, w0, z = 128, 0.02, 500e-9, 0.2, 20e3
    x = (np.arange(N) - N / 2) * pitch
    X, Y = np.meshgrid(x, x)
    E0 = np.exp(-(X ** 2 + Y ** 2) / w0 ** 2)
    assert asm(E0, wvl, pitch, 0) is E0
    E = asm(E0, wvl, pitch, z)
    q = w0 ** 2 + 1j * wvl * z / np.pi            # paraxial Gaussian beam for H = exp(-i pi lambda z f^2)
    ref = w0 ** 2 / q * np.exp(-(X ** 2 + Y ** 2) / q)
    assert np.max(np.abs(E - ref)) < 1e-10
    rs = np.random.RandomState(0)
    F = rs.randn(N, N) + 1j * rs.randn(N, N)
    assert np.isclose(np.sum(np.abs(asm(F, wvl, pitch, 5e3)) ** 2), np.sum(np.abs(F) ** 2), rtol=1e-12)
