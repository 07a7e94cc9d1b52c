"""GPU: the config-2 closed loop (320 px pupil, 80 x 80 subapertures, 72 points, 81 x 81 DM, 2 layers) against
the CPU oracle loop for its first iterations: residual wavefront error, slopes and Strehl traces.

Reference loop: tutorials/AO_closed_loop_Pyramid_WFS.py:601-638 (integrator, gain 0.4, 2-frame delay,
strehlMeter of the image behind the DM). The oracle runs the same loop in float64 from the same initial
atmosphere state and RandomState streams (oracle/atmosphere.LayerOracle), the PWFS oracle on 16 threads
(oracle/pwfs.PyramidOracle), the DM through its pinned separable identity, and the science image through
oracle/science; the reconstructor (a push-pull calibration of the first 1000 KL modes at 10 nm, as the config
specifies) is shared. Tolerances: residual wfe trace 1e-4 relative, per-iteration slopes rel-L2 1e-4, Strehl
1e-3 %. The same loop with the reconstructor in split BF16 (hi + lo on the matrix cores' format, FP32
accumulation) must stay within 1e-4 of the float32 loop (wfe trace and slopes); plain BF16 within 5e-2 (its
stated tolerance: the reconstructor to 2^-8). The factored reconstructor (M2C, M), applied as M2C @ (M @ s)
in float32 and in split BF16, stays within 1e-4 of the float32 loop. Every form's error against the float64
oracle is printed.
"""
import os

import numpy as np
import pytest
from numpy.random import RandomState

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def test_config2_loop_traces_vs_oracle():
    import torch
    from types import SimpleNamespace
    from oopao_amd import Telescope, Source, DeformableMirror, Pyramid, Atmosphere, InteractionMatrix
    from oopao_amd.calibration.compute_KL_modal_basis import compute_KL_basis
    from oopao_amd.closed_loop import run_closed_loop
    from oracle import pwfs as opw, science as osc
    from oracle.atmosphere import LayerOracle, atmosphere_opd
    from oracle.dm import DMOracle
    print('config-2 closed loop vs oracle', flush=True)
    seed = 10
    tel = Telescope(320, 8, 1e-3)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    atm = Atmosphere(tel, 0.15, 25, windSpeed=[10, 20], fractionalR0=[0.7, 0.3], windDirection=[0, 90],
                     altitude=[0, 10000])
    atm.initializeAtmosphere(tel)
    atm.generateNewPhaseScreen(seed)
    dm = DeformableMirror(tel, 80, 0.35, print_dm_properties=False)
    wfs = Pyramid(80, tel, 3, 0.5, n_pix_separation=80, n_pix_edge=40, nTheta_user_defined=72)
    n_modes = int(os.environ.get('OOPAO_CONFIG2_LOOP_MODES', '1000'))
    M2C = compute_KL_basis(tel, SimpleNamespace(r0=0.15, L0=25.0), dm)[:, :n_modes]
    calib = InteractionMatrix(ngs, tel, dm, wfs, M2C, 10e-9, nMeasurements=50, single_pass=False)
    R = M2C @ calib.M
    layers = []
    for i in range(atm.nLayer):
        lay = getattr(atm, 'layer_%d' % (i + 1))
        ms = lay.map.double().cpu().numpy().reshape(lay.M, lay.M)
        rs = RandomState(seed + i * 1000)
        rs.normal(size=lay.B.shape[1])
        ratio = [lay.vX * tel.samplingTime / lay.pixel_size, lay.vY * tel.samplingTime / lay.pixel_size]
        layers.append(LayerOracle(ms, lay.A, lay.B, lay.innerMask, lay.outerMask, rs, ratio))
    # 30 iterations by default (the float64 oracle frame takes ~1 s); longer runs via the environment
    n_iter, gain = int(os.environ.get('OOPAO_CONFIG2_LOOP_ITERS', '30')), 0.4
    forms = {}
    for key, prec, rec in (('fp32', 'fp32', R), ('bf16x3', 'bf16x3', R), ('bf16', 'bf16', R),
                           ('fp32_factored', 'fp32', (M2C, calib.M)), ('bf16x3_factored', 'bf16x3', (M2C, calib.M))):
        atm.generateNewPhaseScreen(seed)            # the same initial screens and RandomState streams
        dm.coefs = np.zeros(dm.nValidAct)
        forms[key] = run_closed_loop(tel, ngs, atm, dm, wfs, rec, gain=gain, n_iter=n_iter, record_slopes=True,
                                     strehl_every=1, reconstructor_precision=prec)
    torch.cuda.synchronize()
    out = forms['fp32']
    flux = opw.flux_map(tel.pupil, ngs.nPhoton, tel.samplingTime, tel.D)
    o = opw.PyramidOracle(320, 80, 3, 0.5, tel.pupil, flux, ngs.wavelength, n_pix_separation=80, n_pix_edge=40,
                          nTheta_user_defined=72, n_jobs=16)
    od = DMOracle(320, 8, 80, 0.35)
    amp = np.sqrt(flux) * tel.pupil
    psf_flat = osc.propagate_field(amp, np.zeros((320, 320)), 2)
    coefs = np.zeros(dm.nValidAct)
    sig = np.zeros(wfs.nSignal)
    wfe, s_trace, sr = [], [], []
    for it in range(n_iter):
        for L in layers:
            L.update()
        opd = (atmosphere_opd(layers, [0.7, 0.3], 320) + od.opd_separable(coefs)) * tel.pupil
        wfe.append(np.std(opd[tel.pupil > 0]) * 1e9)
        sr.append(osc.strehl(osc.propagate_field(amp, opd * 2 * np.pi / o.wavelength, 2), psf_flat))
        s = o.measure_opd(opd)[1]
        s_trace.append(s)
        coefs = coefs - gain * R @ sig
        sig = s
    wfe = np.array(wfe)
    e_wfe = np.max(np.abs(out['wfe_residual'] - wfe) / wfe)
    e_s = max(rel_l2(out['slopes'][k], s_trace[k]) for k in range(n_iter))
    e_sr = np.max(np.abs(out['strehl'] - np.array(sr)))
    print('wfe device', np.round(out['wfe_residual'], 3), 'oracle', np.round(wfe, 3))
    print('Strehl device', np.round(out['strehl'], 4), 'oracle', np.round(sr, 4))
    print('max rel wfe %.2e  worst slopes rel-L2 %.2e  max Strehl diff %.2e %%' % (e_wfe, e_s, e_sr))
    assert e_wfe < 1e-4 and e_s < 1e-4 and e_sr < 1e-3
    for prec, tol in (('bf16x3', 1e-4), ('bf16', 5e-2), ('fp32_factored', 1e-4), ('bf16x3_factored', 1e-4)):
        f = forms[prec]
        d_wfe = np.max(np.abs(f['wfe_residual'] - out['wfe_residual']) / out['wfe_residual'])
        d_s = max(rel_l2(f['slopes'][k], out['slopes'][k]) for k in range(n_iter))
        o_wfe = np.max(np.abs(f['wfe_residual'] - wfe) / wfe)
        o_s = max(rel_l2(f['slopes'][k], s_trace[k]) for k in range(n_iter))
        print('%s reconstructor: vs float32 loop wfe %.2e slopes %.2e; vs float64 oracle wfe %.2e slopes %.2e'
              % (prec, d_wfe, d_s, o_wfe, o_s))
        assert d_wfe < tol and d_s < tol
