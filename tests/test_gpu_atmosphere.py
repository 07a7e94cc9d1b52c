"""GPU: frozen-flow atmosphere update against the float64 oracle.

Reference: Atmosphere.updateLayer / add_row / fill_OPD_support
(Atmosphere.py:410-516, 589-622) and the scikit-image cubic warp
(tools/tools.py:202-209; the oracle's warp is pinned to scikit-image's own
outputs by tests/golden/warp_skimage.npz). Both sides start from the same
mapShift and the same RandomState stream. Tolerance: the device keeps the
screens in float64 and writes the summed OPD in float32: rel-L2 <= 1e-6.
"""
import os

import numpy as np
import pytest
from numpy.random import RandomState

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def _atm(n, Ts, wind, direction, seed):
    from oopao_amd import Telescope, Source, Atmosphere
    tel = Telescope(n, 8, Ts)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    atm = Atmosphere(tel, 0.15, 25, windSpeed=wind, fractionalR0=[0.7, 0.3], windDirection=direction,
                     altitude=[0, 10000])
    atm.initializeAtmosphere(tel)
    atm.generateNewPhaseScreen(seed)
    return tel, ngs, atm


def _oracle_layers(atm, seed):
    from oracle.atmosphere import LayerOracle
    out = []
    for i in range(atm.nLayer):
        lay = getattr(atm, 'layer_%d' % (i + 1))
        ms = lay.map.double().cpu().numpy().reshape(lay.M, lay.M)
        rs = RandomState(seed + i * 1000)
        rs.normal(size=lay.B.shape[1])          # the draw made by generateNewPhaseScreen
        ratio = [lay.vX * atm.telescope.samplingTime / lay.pixel_size,
                 lay.vY * atm.telescope.samplingTime / lay.pixel_size]
        out.append(LayerOracle(ms, lay.A, lay.B, lay.innerMask, lay.outerMask, rs, ratio))
    return out


@pytest.mark.parametrize('wind,direction,Ts', [([10, 20], [0, 90], 0.02),      # 0.8 px and 1.6 px per step
                                               ([15, 25], [45, 200], 0.01),    # diagonal, negative components
                                               ([10, 20], [0, 90], 0.001)])    # config-2 speeds: 0.04 / 0.08 px
def test_update_matches_oracle(wind, direction, Ts):
    from oracle.atmosphere import atmosphere_opd
    seed = 10
    tel, ngs, atm = _atm(32, Ts, wind, direction, seed)
    layers = _oracle_layers(atm, seed)
    ref0 = atmosphere_opd(layers, atm.fractionalR0, 32)
    assert rel_l2(atm.OPD, ref0) < 1e-6
    worst = 0.0
    steps = int(os.environ.get('OOPAO_ATM_STEPS', '30'))      # longer runs via the environment
    for step in range(steps):
        atm.update()
        for L in layers:
            L.update()
        err = rel_l2(atm.OPD, atmosphere_opd(layers, atm.fractionalR0, 32))
        worst = max(worst, err)
        # the clip range the warp uses: the extruded map's exact min / max (folded into the extrusion launch)
        for i in range(atm.nLayer):
            lay = getattr(atm, 'layer_%d' % (i + 1))
            mm = lay.mm.cpu().numpy()
            assert (mm[0], mm[1]) == (float(lay.map.min()), float(lay.map.max())), (step, i, mm[:2])
    print('worst rel_l2 over %d steps' % steps, worst)
    assert worst < 1e-6


def test_paired_telescope_propagates_atmosphere():
    tel, ngs, atm = _atm(32, 0.01, [10, 20], [0, 90], 3)
    tel + atm
    atm.update()
    ref = atm.OPD * tel.pupil
    assert rel_l2(tel.OPD, ref) < 1e-7
    ngs ** atm * tel
    assert rel_l2(ngs.OPD, ref) < 1e-7


def test_property_setters():
    # elevation / L0 / fractionalR0 / windSpeed setters (Atmosphere.py:1042-1161)
    tel, ngs, atm = _atm(32, 0.01, [10, 20], [0, 90], 3)
    r0_zenith = atm.r0_zenith
    assert abs(r0_zenith - 0.15) < 1e-15                      # elevation 90: slant = zenith
    # elevation 30 deg: slant distances h_z / sin(el), r0 = r0_zenith sin(el)^(3/5), layers rebuilt
    atm.elevation = 30
    el = np.radians(30.0)
    assert np.allclose(atm.altitude, [0.0, 10000 / np.sin(el)], rtol=1e-14)
    assert abs(atm.r0 - r0_zenith * np.sin(el) ** 0.6) < 1e-15
    assert atm.layer_2.altitude == atm.altitude[1]
    assert np.isfinite(atm.OPD).all() and atm.OPD.std() > 0
    atm.elevation = 2                                        # clamped to 5 deg
    assert atm.elevation == 5.0
    # L0: the layers are rebuilt with covariances of the new outer scale
    xxt_25 = atm.layer_1.XXt.copy()
    atm.L0 = 10
    assert atm.L0 == 10
    assert not np.allclose(atm.layer_1.XXt, xxt_25)
    from oopao_amd.phaseStats import makeCovarianceMatrix
    lay = atm.layer_1
    assert np.allclose(atm.layer_1.XXt, makeCovarianceMatrix(lay.outerZ, lay.outerZ, 10, atm.r0_def), rtol=1e-12)
    # fractionalR0 / windSpeed: the setters' V0 = sum(fr * v) (Atmosphere.py:1067, 1131)
    atm.fractionalR0 = [0.5, 0.5]
    assert abs(atm.V0 - (0.5 * 10 + 0.5 * 20)) < 1e-12
    atm.windSpeed = [5, 15]
    assert abs(atm.V0 - (0.5 * 5 + 0.5 * 15)) < 1e-12
    assert abs(atm.tau0 - 0.31 * atm.r0 / atm.V0) < 1e-15
    from oopao_amd.tools import OopaoError
    with pytest.raises(OopaoError):
        atm.fractionalR0 = [1.0]


def test_save_load_state_resumes_the_same_screens(tmp_path):
    # a checkpoint taken mid-run resumes bit for bit, on the same object and on a second atmosphere built
    # with the same parameters (0.8 / 1.6 px per step: whole-pixel extrusions and the sub-pixel buffer)
    tel, ngs, atm = _atm(48, 0.02, [10, 20], [30, 120], 7)
    for _ in range(5):
        atm.update()
    path = str(tmp_path / 'atm_state.npz')
    atm.save_state(path)
    ref = []
    for _ in range(7):
        atm.update()
        ref.append(atm._OPD_dev.clone())
    atm.load_state(path)
    for k in range(7):
        atm.update()
        assert torch_equal(atm._OPD_dev, ref[k]), k
    tel2, ngs2, atm2 = _atm(48, 0.02, [10, 20], [30, 120], 99)
    atm2.load_state(path)
    for k in range(7):
        atm2.update()
        assert torch_equal(atm2._OPD_dev, ref[k]), k
    tel3, ngs3, atm3 = _atm(40, 0.02, [10, 20], [30, 120], 7)
    from oopao_amd.tools import OopaoError
    with pytest.raises(OopaoError):
        atm3.load_state(path)


def torch_equal(a, b):
    import torch
    return bool(torch.equal(a, b))
