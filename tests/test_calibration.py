"""Interaction matrix, CalibrationVault and the closed loop.

CPU: CalibrationVault (calibration/CalibrationVault.py:45-70) checked against
the Moore-Penrose identities and numpy's float64 pinv.
GPU: InteractionMatrix (calibration/InteractionMatrix.py:13-169) against the
oracle's columns; the device closed loop against a CPU oracle loop
(tutorials/AO_closed_loop_Pyramid_WFS.py:601-638) driven by the same
reconstructor and the same atmosphere stream.
"""
import numpy as np
import pytest
from numpy.random import RandomState

from tests.helpers import rel_l2


def test_calibration_vault_pinv_identities():
    from oopao_amd.calibration.CalibrationVault import CalibrationVault
    rs = np.random.RandomState(0)
    D = rs.randn(60, 25) @ np.diag(np.logspace(0, -3, 25))
    cv = CalibrationVault(D, device='cpu')
    M = cv.M
    assert np.allclose(D @ M @ D, D, atol=1e-10)
    assert np.allclose(M @ D @ M, M, atol=1e-6)
    assert np.allclose(M, np.linalg.pinv(D), rtol=1e-8, atol=1e-8)
    s = np.linalg.svd(D, compute_uv=False)
    assert np.allclose(cv.eigenValues, s, rtol=1e-10)
    assert np.isclose(cv.cond, s[0] / s[-1], rtol=1e-10)
    cv.nTrunc = 5
    assert cv.Mtrunc.shape == (25, 60)
    assert np.isclose(cv.cond, cv.eigenValues[0] / cv.eigenValues[-6])


@pytest.mark.parametrize('shape', [(300, 500), (700, 300)])
def test_calibration_vault_gram_route(shape):
    # the Gram route ('auto' for min side >= 256 and cond <= 1e4): singular values, pseudo-inverse and
    # factor orthonormality against numpy's float64 SVD at cond 1e3 (errors ~eps cond^2 on M)
    from oopao_amd.calibration.CalibrationVault import CalibrationVault
    m, k = shape
    r = min(m, k)
    rs = np.random.RandomState(1)
    Ua = np.linalg.qr(rs.randn(m, r))[0]
    Va = np.linalg.qr(rs.randn(k, r))[0]
    D = Ua @ np.diag(np.logspace(0, -3, r)) @ Va.T
    cv = CalibrationVault(D, device='cpu')
    s = np.linalg.svd(D, compute_uv=False)
    assert np.allclose(cv.eigenValues, s, rtol=1e-9)
    pinv = np.linalg.pinv(D)
    assert np.abs(cv.M - pinv).max() <= 1e-8 * np.abs(pinv).max()
    assert np.allclose(cv.U @ cv.U.T, np.eye(r), atol=1e-9)
    assert np.allclose(cv.V @ cv.V.T, np.eye(r), atol=1e-6)
    assert np.allclose(cv.D, D, atol=1e-12)
    assert np.isclose(cv.cond, 1e3, rtol=1e-8)
    # an ill-conditioned matrix takes the SVD route (same results as method='svd')
    D2 = Ua @ np.diag(np.logspace(0, -6, r)) @ Va.T
    a = CalibrationVault(D2, device='cpu')
    b = CalibrationVault(D2, device='cpu', method='svd')
    assert np.allclose(a.eigenValues, b.eigenValues, rtol=1e-12)
    with pytest.raises(ValueError):
        CalibrationVault(D2, device='cpu', method='gram')


def _small_loop_system(seed=10):
    from oopao_amd import Telescope, Source, DeformableMirror, Pyramid, Atmosphere
    tel = Telescope(32, 8, 0.005)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    atm = Atmosphere(tel, 0.15, 25, windSpeed=[10, 20], fractionalR0=[0.7, 0.3], windDirection=[0, 90],
                     altitude=[0, 10000])
    atm.initializeAtmosphere(tel)
    atm.generateNewPhaseScreen(seed)
    dm = DeformableMirror(tel, 8, 0.35, print_dm_properties=False)
    wfs = Pyramid(8, tel, 2, 0.1, n_pix_separation=8, n_pix_edge=4, nTheta_user_defined=16)
    return tel, ngs, atm, dm, wfs


def _oracle_pwfs(wfs, tel, ngs):
    from oracle.pwfs import PyramidOracle, flux_map
    flux = flux_map(tel.pupil, ngs.nPhoton, tel.samplingTime, tel.D)
    return PyramidOracle(tel.resolution, wfs.nSubap, wfs.modulation, wfs.lightRatio, tel.pupil, flux, ngs.wavelength,
                         n_pix_separation=wfs.n_pix_separation, n_pix_edge=wfs.n_pix_edge,
                         nTheta_user_defined=wfs.nTheta)


@pytest.mark.gpu
def test_interaction_matrix_matches_oracle():
    from oopao_amd import InteractionMatrix
    from oracle.dm import DMOracle
    tel, ngs, atm, dm, wfs = _small_loop_system()
    o = _oracle_pwfs(wfs, tel, ngs)
    assert np.array_equal(o.validI4Q, wfs.validI4Q)
    od = DMOracle(32, 8, 8, 0.35)
    modes = od.modes()
    M2C = np.eye(dm.nValidAct)[:, :24]
    stroke = 20e-9
    calib = InteractionMatrix(ngs, tel, dm, wfs, M2C, stroke, nMeasurements=10, single_pass=True)
    D = calib.D
    assert D.shape == (wfs.nSignal, 24)
    Dref = np.stack([o.measure_opd(od.opd(modes, M2C[:, k] * stroke))[1] / stroke for k in range(24)], axis=1)
    err = rel_l2(D, Dref)
    print('IM rel_l2 vs fp64 oracle', err)
    assert err < 1e-4
    calib_pp = InteractionMatrix(ngs, tel, dm, wfs, M2C, stroke, nMeasurements=24, single_pass=False)
    Dpp = np.stack([0.5 * (o.measure_opd(od.opd(modes, M2C[:, k] * stroke))[1]
                           - o.measure_opd(od.opd(modes, -M2C[:, k] * stroke))[1]) / stroke for k in range(24)], axis=1)
    assert rel_l2(calib_pp.D, Dpp) < 1e-4
    # the calibration measures slopes-only frames; read afterwards, the camera frame is the whole frame of
    # the last measurement (the pulls of the last batch)
    last = np.asarray(wfs.cam.frame)            # a batch reads as [res, res, B], as OOPAO's cubes
    assert last.shape == (wfs.cam.resolution, wfs.cam.resolution, 24)
    ph = od.opd(modes, -M2C[:, 23] * stroke) * tel.pupil * 2 * np.pi / ngs.wavelength
    assert rel_l2(last[:, :, 23], o.frame(ph)) < 1e-5


@pytest.mark.gpu
def test_interaction_matrix_from_phase_screen_matches_oracle():
    # calibration/InteractionMatrix.py:172-259: push-pull columns of OPD maps [n, n, nModes], with and
    # without a phase offset added to both maps, batch sizes that leave a ragged last batch
    from oopao_amd import InteractionMatrixFromPhaseScreen
    from oracle.dm import DMOracle
    tel, ngs, atm, dm, wfs = _small_loop_system()
    o = _oracle_pwfs(wfs, tel, ngs)
    od = DMOracle(32, 8, 8, 0.35)
    modes = od.modes()
    rs = np.random.RandomState(4)
    screens = np.stack([od.opd(modes, rs.randn(dm.nValidAct)) for _ in range(7)], axis=2)
    screens /= np.abs(screens).max()
    stroke = 15e-9
    offset = 0.3 * rs.randn(32, 32)
    k = ngs.wavelength / (2 * np.pi)
    for off, nMeas in ((0, 3), (offset, 7), (offset, 1)):
        calib = InteractionMatrixFromPhaseScreen(ngs, atm, tel, wfs, screens, stroke, phaseOffset=off,
                                                 nMeasurements=nMeas, invert=False, print_time=False)
        base = off * k if np.ndim(off) == 2 else 0
        Dref = np.stack([0.5 * (o.measure_opd(screens[:, :, j] * stroke + base)[1]
                                - o.measure_opd(-screens[:, :, j] * stroke + base)[1]) / stroke
                         for j in range(7)], axis=1)
        assert calib.D.shape == (wfs.nSignal, 7)
        err = rel_l2(calib.D, Dref)
        print('IM from phase screens, nMeasurements', nMeas, 'rel_l2', err)
        assert err < 1e-4
    one = InteractionMatrixFromPhaseScreen(ngs, atm, tel, wfs, screens[:, :, 2], stroke, invert=False,
                                           print_time=False)
    assert one.D.shape == (wfs.nSignal, 1)
    ref_one = 0.5 * (o.measure_opd(screens[:, :, 2] * stroke)[1] - o.measure_opd(-screens[:, :, 2] * stroke)[1]) / stroke
    assert rel_l2(one.D[:, 0], ref_one) < 1e-4


def _oracle_layers(seed, tel, atm):
    """Oracle layers from the device atmosphere's current state and the same RandomState streams."""
    from oracle.atmosphere import LayerOracle
    layers = []
    for i in range(atm.nLayer):
        lay = getattr(atm, 'layer_%d' % (i + 1))
        ms = lay.map.double().cpu().numpy().reshape(lay.M, lay.M)
        rs = RandomState(seed + i * 1000)
        rs.normal(size=lay.B.shape[1])
        ratio = [lay.vX * tel.samplingTime / lay.pixel_size, lay.vY * tel.samplingTime / lay.pixel_size]
        layers.append(LayerOracle(ms, lay.A, lay.B, lay.innerMask, lay.outerMask, rs, ratio))
    return layers


def _oracle_loop(layers, tel, ngs, dm, wfs, R, n_iter, gain, strehl=False):
    """The closed loop of tutorials/AO_closed_loop_Pyramid_WFS.py:601-638 on the CPU oracle."""
    from oracle.dm import DMOracle
    from oracle.atmosphere import atmosphere_opd
    from oracle import science as osc
    o = _oracle_pwfs(wfs, tel, ngs)
    amp = np.sqrt(o.flux) * tel.pupil
    psf_flat = osc.propagate_field(amp, np.zeros((32, 32)), 2) if strehl else None
    od = DMOracle(32, 8, 8, 0.35)
    modes = od.modes()
    coefs = np.zeros(dm.nValidAct)
    sig = np.zeros(wfs.nSignal)
    wfe, s_trace, sr_trace = [], [], []
    for it in range(n_iter):
        for L in layers:
            L.update()
        opd = (atmosphere_opd(layers, [0.7, 0.3], 32) + od.opd(modes, coefs)) * tel.pupil
        wfe.append(np.std(opd[tel.pupil > 0]) * 1e9)
        if strehl:
            sr_trace.append(osc.strehl(osc.propagate_field(amp, opd * 2 * np.pi / o.wavelength, 2), psf_flat))
        s = o.measure_opd(opd)[1]
        s_trace.append(s)
        coefs = coefs - gain * R @ sig
        sig = s
    return np.array(wfe), s_trace, sr_trace


@pytest.mark.gpu
def test_closed_loop_matches_oracle_loop():
    from oopao_amd import InteractionMatrix
    from oopao_amd.closed_loop import run_closed_loop
    seed = 10
    tel, ngs, atm, dm, wfs = _small_loop_system(seed)
    calib = InteractionMatrix(ngs, tel, dm, wfs, np.eye(dm.nValidAct), 20e-9, nMeasurements=100)
    calib.nTrunc = 10
    R = calib.Mtrunc
    n_iter, gain = 25, 0.4
    layers = _oracle_layers(seed, tel, atm)
    out = run_closed_loop(tel, ngs, atm, dm, wfs, R, gain=gain, n_iter=n_iter, record_slopes=True, strehl_every=1)
    wfe, s_trace, sr_trace = _oracle_loop(layers, tel, ngs, dm, wfs, R, n_iter, gain, strehl=True)
    print('residual wfe (nm) device', out['wfe_residual'][-5:], 'oracle', wfe[-5:])
    assert np.max(np.abs(out['wfe_residual'] - wfe) / wfe) < 1e-4
    err = max(rel_l2(out['slopes'][k], s_trace[k]) for k in range(n_iter))
    print('worst per-iteration slope rel_l2', err)
    assert err < 1e-4
    assert out['wfe_residual'][-1] < 0.5 * out['wfe_atmosphere'][-1]      # the loop corrects
    # Strehl trace (strehlMeter on the image behind the DM, tools/tools.py:102-157)
    sr_err = np.max(np.abs(out['strehl'] - np.array(sr_trace)))
    print('Strehl trace device', out['strehl'][-3:], 'oracle', sr_trace[-3:], 'max abs diff [%]', sr_err)
    assert sr_err < 1e-3


@pytest.mark.gpu
def test_closed_loop_bf16_reconstructor():
    """BF16 reconstructor (FP32 accumulation) against the oracle loop driven by the BF16-rounded
    reconstructor (same tolerances as the float32 loop), and the stated distance of that loop from the
    float64-reconstructor loop."""
    from oopao_amd import InteractionMatrix
    from oopao_amd.closed_loop import run_closed_loop
    from oracle.contract import bf16_rne
    seed = 11
    tel, ngs, atm, dm, wfs = _small_loop_system(seed)
    calib = InteractionMatrix(ngs, tel, dm, wfs, np.eye(dm.nValidAct), 20e-9, nMeasurements=100)
    calib.nTrunc = 10
    R = calib.Mtrunc
    R16 = bf16_rne(R.astype(np.float32)).astype(np.float64)
    print('reconstructor BF16 rel-L2', rel_l2(R16, R))
    n_iter, gain = 25, 0.4
    layers = _oracle_layers(seed, tel, atm)
    out = run_closed_loop(tel, ngs, atm, dm, wfs, R, gain=gain, n_iter=n_iter, record_slopes=True,
                          reconstructor_precision='bf16')
    wfe16, s16, _ = _oracle_loop(layers, tel, ngs, dm, wfs, R16, n_iter, gain)
    print('residual wfe (nm) device bf16', out['wfe_residual'][-5:], 'oracle bf16', wfe16[-5:])
    assert np.max(np.abs(out['wfe_residual'] - wfe16) / wfe16) < 1e-4
    assert max(rel_l2(out['slopes'][k], s16[k]) for k in range(n_iter)) < 1e-4
    assert out['wfe_residual'][-1] < 0.5 * out['wfe_atmosphere'][-1]
