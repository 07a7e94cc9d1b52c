"""GPU: boiling layers (Atmosphere.apply_boiling, Atmosphere.py:387-408) against the float64 oracle.

Each step a boiling layer's extended screen becomes alpha mapShift + beta phi, phi a fresh phase screen
(ft_sh_phase_screen in mode 2, ft_phase_screen otherwise, phaseStats.py:153-271) of seed boiling_seed + 1,
before the frozen-flow shift. The device builds phi from the host RandomState draws with its own
centred-DFT GEMMs and subharmonic kernel; the oracle with numpy FFTs. Both start from the same
mapShift and RandomState streams. Tolerance: rel-L2 <= 1e-6 on the summed float32 OPD at every step.
Also: checkpoint / resume with boiling, and load_state refusing a state taken with other parameters.
"""
import numpy as np
import pytest
from numpy.random import RandomState

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def _atm(n, Ts, wind, t_boiling, mode=2, seed=10):
    from oopao_amd import Telescope, Source, Atmosphere
    tel = Telescope(n, 8, Ts)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    atm = Atmosphere(tel, 0.15, 25, windSpeed=wind, fractionalR0=[0.7, 0.3], windDirection=[0, 90],
                     altitude=[0, 10000], mode=mode, t_boiling=t_boiling)
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
        Ts = atm.telescope.samplingTime
        L = LayerOracle(ms, lay.A, lay.B, lay.innerMask, lay.outerMask, rs,
                        [lay.vX * Ts / lay.pixel_size, lay.vY * Ts / lay.pixel_size])
        assert lay.boiling_seed == 100000 + seed + i * 100000          # Atmosphere.py:797-799
        L.set_boiling(atm.t_boiling[i], Ts, atm.r0, atm.L0, lay.D / lay.resolution, lay.boiling_seed, atm.mode)
        out.append(L)
    return out


@pytest.mark.parametrize('wind,t_boiling,Ts,mode', [([10, 0], [0.01, 0.004], 1e-3, 2),   # windless boiling layer
                                                    ([10, 20], 0.05, 0.02, 2),          # 0.8 / 1.6 px per step
                                                    ([10, 20], [None, 0.003], 1e-3, 1)])  # ft_phase_screen
def test_boiling_matches_oracle(wind, t_boiling, Ts, mode):
    from oracle.atmosphere import atmosphere_opd
    seed = 10
    n = 80
    print('boiling: wind %s t_boiling %s Ts %g mode %d' % (wind, t_boiling, Ts, mode), flush=True)
    tel, ngs, atm = _atm(n, Ts, wind, t_boiling, mode, seed)
    layers = _oracle_layers(atm, seed)
    assert rel_l2(atm.OPD, atmosphere_opd(layers, atm.fractionalR0, n)) < 1e-6
    worst = 0.0
    start = atm.OPD.copy()
    for _ in range(25):
        atm.update()
        for L in layers:
            L.update()
        worst = max(worst, rel_l2(atm.OPD, atmosphere_opd(layers, atm.fractionalR0, n)))
    print('worst rel_l2 over 25 steps %.3e' % worst)
    assert worst < 1e-6
    for i, L in enumerate(layers):
        assert getattr(atm, 'layer_%d' % (i + 1)).boiling_seed == L.boiling_seed
    # the screens did evolve (boiling on every layer with a boiling time)
    assert rel_l2(atm.OPD, start) > 0.05


def test_boiling_checkpoint_resume_and_parameter_checks(tmp_path):
    from oopao_amd.tools import OopaoError
    tel, ngs, atm = _atm(48, 1e-3, [10, 0], [0.01, 0.004])
    for _ in range(4):
        atm.update()
    path = str(tmp_path / 'atm.npz')
    atm.save_state(path)
    ref = []
    for _ in range(4):
        atm.update()
        ref.append(atm.OPD.copy())
    tel2, ngs2, atm2 = _atm(48, 1e-3, [10, 0], [0.01, 0.004], seed=3)
    atm2.load_state(path)
    for k in range(4):
        atm2.update()
        assert np.array_equal(atm2.OPD, ref[k])
    # another wind, boiling time or sampling time: refused
    for kw in (dict(wind=[12, 0], t_boiling=[0.01, 0.004], Ts=1e-3), dict(wind=[10, 0], t_boiling=[0.01, 0.005], Ts=1e-3),
               dict(wind=[10, 0], t_boiling=[0.01, 0.004], Ts=2e-3)):
        _, _, other = _atm(48, kw['Ts'], kw['wind'], kw['t_boiling'])
        with pytest.raises(OopaoError):
            other.load_state(path)


def test_t_boiling_setter():
    from oopao_amd.tools import OopaoError
    tel, ngs, atm = _atm(48, 1e-3, [10, 20], None)
    assert atm.t_boiling == [None, None]
    atm.t_boiling = 0.02
    assert atm.layer_1.t_boiling == 0.02 and atm.layer_2.t_boiling == 0.02
    with pytest.raises(OopaoError):
        atm.t_boiling = [0.1, 0.2, 0.3]
    s0 = atm.layer_1.boiling_seed
    atm.update()
    assert atm.layer_1.boiling_seed == s0 + 1
    atm.t_boiling = 1e4                     # alpha >= 1 - 1e-6: frozen, no new screen drawn
    atm.update()
    assert atm.layer_1.boiling_seed == s0 + 1
