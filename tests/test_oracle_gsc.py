"""CPU: pins of the gain-sensing-camera oracle (oracle/gsc.py)."""
import numpy as np

from oracle import gsc


def _dft(N, sign):
    k = np.arange(N) - N // 2
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / N)


def test_centred_transforms_match_explicit_dft():
    rs = np.random.RandomState(0)
    for N in (6, 8, 12):
        x = rs.randn(N, N) + 1j * rs.randn(N, N)
        D, Di = _dft(N, -1), _dft(N, +1)
        assert np.allclose(gsc.centred_fft2(x), D @ x @ D.T / N, atol=1e-12)
        assert np.allclose(gsc.centred_ifft2(x), Di @ x @ Di.T / N, atol=1e-12)
        assert np.allclose(gsc.centred_ifft2(gsc.centred_fft2(x)), x, atol=1e-12)


def test_basis_product_real_and_parseval():
    rs = np.random.RandomState(1)
    B = rs.randn(10, 10, 4)
    bp = gsc.basis_product(B, 16)
    F = gsc.centred_fft2(gsc.pad_basis(B, 16))
    assert np.max(np.abs(bp.imag)) < 1e-12 * np.max(np.abs(bp))
    assert np.allclose(bp.real, np.abs(F) ** 2, rtol=1e-12, atol=1e-14)
    assert np.allclose(bp.real.sum(axis=(0, 1)), (B ** 2).sum(axis=(0, 1)), rtol=1e-12)


def test_impulse_response_flat_mask_is_zero():
    rs = np.random.RandomState(2)
    frame = rs.rand(16, 16)
    frame /= frame.sum()
    ir = gsc.impulse_response(np.ones((16, 16), dtype=complex), frame)
    assert np.max(np.abs(ir)) < 1e-15


def test_optical_gains_identity_and_scale():
    rs = np.random.RandomState(3)
    N = 16
    x = np.arange(N) - N / 2
    X, Y = np.meshgrid(x, x)
    mask = np.exp(1j * 0.7 * (np.abs(X) + np.abs(Y)))          # pyramid-like phase mask
    basis = rs.randn(8, 8, 5)
    f0 = rs.rand(N, N)
    f1 = rs.rand(N, N)
    o = gsc.GSCOracle(mask, basis)
    o.calibration(f0)
    assert np.allclose(o.optical_gains(f0), 1.0, atol=1e-12)
    assert np.allclose(o.optical_gains(3.7 * f1), o.optical_gains(f1), rtol=1e-12)
