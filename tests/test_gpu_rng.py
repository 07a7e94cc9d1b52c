"""GPU: the atmosphere's device RNG mode (Atmosphere(..., rng='device')).

* oopao_normal_philox against the CPU restatement (oracle/rng.py, pinned by the Philox known-answer
  vectors): the same normals to 1e-12 (double-precision log / sin / cos of two libraries).
* A two-layer atmosphere (one layer windless and boiling, one windy and boiling) over 25 steps against the
  float64 oracle layers fed with the same counter-based streams: extrusion draws from key
  (seed + 1000 i, 0), boiling screens from key (boiling_seed, 1) -- rel-L2 <= 1e-6 on the summed OPD.
* Checkpoint / resume in device mode is bit for bit; a state saved in the other mode is refused.
"""
import ctypes

import numpy as np
import pytest
import torch
from numpy.random import RandomState

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('count,key0,key1,stream,draw', [(1, 7, 0, 0, 0), (1001, 1042, 0, 0, 5),
                                                         (1 << 20, 100011, 1, 0, 0), (33, 5, 0, 3, (1 << 32) + 9)])
def test_device_normals_equal_restatement(count, key0, key1, stream, draw):
    from oopao_amd import _lib
    from oracle import rng as orng
    dev = torch.device('cuda', 0)
    out = torch.empty(count, dtype=torch.float64, device=dev)
    _lib.call('oopao_normal_philox', ctypes.c_void_p(out.data_ptr()), count, key0, key1, stream, draw,
              ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
    ref = orng.normals(count, key0, key1, stream, draw)
    assert np.max(np.abs(out.cpu().numpy() - ref)) <= 1e-12


def _atm(n, Ts, wind, t_boiling, rng, seed=10):
    from oopao_amd import Telescope, Source, Atmosphere
    tel = Telescope(n, 8, Ts)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    atm = Atmosphere(tel, 0.15, 25, windSpeed=wind, fractionalR0=[0.7, 0.3], windDirection=[0, 90],
                     altitude=[0, 10000], t_boiling=t_boiling, rng=rng)
    atm.initializeAtmosphere(tel)
    atm.generateNewPhaseScreen(seed)
    return tel, ngs, atm


def test_device_rng_atmosphere_matches_oracle():
    from oracle import rng as orng
    from oracle.atmosphere import LayerOracle, atmosphere_opd
    seed, n, Ts = 10, 80, 0.02
    print('device RNG atmosphere: wind [10, 20] (0.8 / 1.6 px per step), boiling', flush=True)
    tel, ngs, atm = _atm(n, Ts, [0, 20], [0.05, 0.08], 'device', seed)
    layers = []
    for i in range(atm.nLayer):
        lay = getattr(atm, 'layer_%d' % (i + 1))
        ms = lay.map.double().cpu().numpy().reshape(lay.M, lay.M)
        assert lay.rng_key == seed + i * 1000 and lay.rng_draw == 0
        L = LayerOracle(ms, lay.A, lay.B, lay.innerMask, lay.outerMask, orng.PhiloxStream(seed + i * 1000, 0, 0),
                        [lay.vX * Ts / lay.pixel_size, lay.vY * Ts / lay.pixel_size])
        L.set_boiling(atm.t_boiling[i], Ts, atm.r0, atm.L0, lay.D / lay.resolution, lay.boiling_seed, atm.mode,
                      draws=lambda s, c: orng.normals(c, s & 0xFFFFFFFF, 1, 0, 0))
        layers.append(L)
    worst = 0.0
    for _ in range(25):
        atm.update()
        for L in layers:
            L.update()
        worst = max(worst, rel_l2(atm.OPD, atmosphere_opd(layers, atm.fractionalR0, n)))
    print('device RNG: worst rel_l2 over 25 steps %.3e, extrusion draws %d' % (worst, atm.layer_2.rng_draw))
    assert worst < 1e-6
    assert atm.layer_2.rng_draw == layers[1].rs.draw > 0


def test_device_rng_checkpoint(tmp_path):
    from oopao_amd.tools import OopaoError
    tel, ngs, atm = _atm(48, 0.02, [10, 20], [None, 0.05], 'device')
    for _ in range(3):
        atm.update()
    path = str(tmp_path / 'atm.npz')
    atm.save_state(path)
    ref = []
    for _ in range(3):
        atm.update()
        ref.append(atm.OPD.copy())
    tel2, ngs2, atm2 = _atm(48, 0.02, [10, 20], [None, 0.05], 'device', seed=4)
    atm2.load_state(path)
    for k in range(3):
        atm2.update()
        assert np.array_equal(atm2.OPD, ref[k])
    _, _, host = _atm(48, 0.02, [10, 20], [None, 0.05], 'host')
    with pytest.raises(OopaoError):
        host.load_state(path)
