"""GPU: Detector.integrate / readout chain (Detector.py:192-354): binning (applied twice, as the
reference does), ADC quantisation and unsigned output, saturation, dark current, EMCCD gain, log
scale. Deterministic stages are compared exactly with numpy; noise stages by their moments."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _frame(n=64, seed=0):
    return np.random.RandomState(seed).uniform(0, 100, (n, n))


def _run(det, f):
    det.integrate(torch.as_tensor(f, dtype=torch.float32, device=det.device))
    return det.frame


def test_binning_applied_twice():
    from oopao_amd import Detector
    f = _frame()
    out = _run(Detector(64, binning=2), f)
    ref = f.reshape(16, 4, 16, 4).sum(axis=(1, 3))
    assert out.shape == (16, 16)
    assert np.allclose(out, ref, rtol=1e-5)


def test_adc_quantisation_and_output_type():
    from oopao_amd import Detector
    f = _frame(seed=1) * 15.0                    # up to 1500 e-: saturates a 1000 e- well
    det = Detector(64, bits=8, FWC=1000)
    out = _run(det, f)
    g = np.clip(f.astype(np.float32), 0, 1000) / 1000 * 255
    ref = np.clip(g, max(g.min(), 0), 255).astype(np.uint8)
    assert out.dtype == np.uint8
    assert np.abs(out.astype(int) - ref.astype(int)).max() <= 1   # float32 rounding at integer edges
    assert np.mean(out != ref) < 1e-3
    assert det.saturation > 100
    assert abs(det.quantification_noise - 1000 * 2 ** -8 / np.sqrt(12)) < 1e-12


def test_dark_current_and_emccd_moments():
    from oopao_amd import Detector
    f = np.full((256, 256), 50.0)
    det = Detector(256, integrationTime=2.0, darkCurrent=25.0)
    det.seed_noise(3)
    det._integrated_time = 2.0
    out = _run(det, f)
    assert abs(out.mean() - 100.0) < 0.2         # 50 + Poisson(25 * 2): sd of the mean 0.03
    assert abs(out.var() - 50.0) < 2.0
    em = Detector(256, sensor='EMCCD', gain=10)
    em.seed_noise(4)
    out = _run(em, f)
    assert abs(out.mean() / 500.0 - 1) < 0.01
    assert abs(out.var() / (100 * 50.0) - 1) < 0.05


def test_log_scale():
    from oopao_amd import Detector
    f = _frame(seed=2) + 1.0
    out = _run(Detector(64, log_scale=True), f)
    assert np.allclose(out, np.log10(f.astype(np.float32)), rtol=1e-6)
