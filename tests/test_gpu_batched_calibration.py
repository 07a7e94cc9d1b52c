"""GPU: push-pull calibration of several PWFS systems recorded into one HIP graph
(oopao_amd/calibration/batched.py) against the per-system InteractionMatrix
(calibration/InteractionMatrix.py:13-169 semantics: 0.5 (s_push - s_pull) / stroke).

Tolerances: graph replays are bit-identical to the eager run of the same recorded work
(checked on three replays with eager work and synchronisations between them); against
InteractionMatrix, whose tile settings and float64 push-pull combination differ only in
summation order, rel-L2 <= 1e-5 per system.
"""
import numpy as np
import pytest
import torch

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def _system(n, nSubap, T, seed, n_modes):
    from oopao_amd import Telescope, Source, DeformableMirror, Pyramid
    tel = Telescope(n, 8, 1e-3)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    dm = DeformableMirror(tel, nSubap, 0.35, print_dm_properties=False)
    wfs = Pyramid(nSubap, tel, 2, 0.1, n_pix_separation=nSubap, n_pix_edge=nSubap // 2, nTheta_user_defined=T)
    M2C = np.random.RandomState(seed).randn(dm.nValidAct, n_modes)
    return ngs, tel, dm, wfs, M2C


def test_graph_calibration_two_systems():
    from oopao_amd import InteractionMatrix
    from oopao_amd.calibration.batched import BatchedCalibration
    print('systems n=32 (T=16) and n=64 (T=24)', flush=True)
    systems = [_system(32, 8, 16, 1, 37), _system(64, 16, 24, 2, 50)]
    stroke = 10e-9
    bc = BatchedCalibration(systems, stroke, tune=True)
    print('autotuner', bc.tuning, flush=True)
    for t in bc.tuning:
        assert t['batch'] >= 1 and t['nchunk'] >= 1 and t['pts_per_task'] >= 1
    eager = [D.clone() for D in bc.run_eager()]
    torch.cuda.synchronize()
    bc.capture()
    for rep in range(3):
        for cs in bc.systems:
            cs.D.fill_(float('nan'))
        Ds = bc.run()
        torch.cuda.synchronize()
        for D, E in zip(Ds, eager):
            assert torch.equal(D, E), 'replay %d differs from the eager run' % rep
        # unrelated eager work between replays
        x = torch.randn(1000, device='cuda')
        (x * 2).sum().item()
    for (ngs, tel, dm, wfs, M2C), D in zip(systems, eager):
        ref = InteractionMatrix(ngs, tel, dm, wfs, M2C, stroke, nMeasurements=16, single_pass=False, invert=False)
        err = rel_l2(D.double().cpu().numpy(), ref.D_dev.cpu().numpy())
        print('n=%d rel-L2 vs InteractionMatrix %.2e' % (tel.resolution, err))
        assert err < 1e-5


def test_autotuner_cache_roundtrip(tmp_path):
    from oopao_amd.calibration.batched import PWFSAutotuner, CalibrationSystem
    print('autotuner cache', flush=True)
    ngs, tel, dm, wfs, M2C = _system(32, 8, 16, 3, 8)
    tuner = PWFSAutotuner(batch_candidates=(4, 8))
    cs = CalibrationSystem(ngs, tel, dm, wfs, M2C, 10e-9)
    best = cs.tune(tuner)
    path = tmp_path / 'tune.json'
    tuner.save(str(path))
    t2 = PWFSAutotuner()
    t2.load(str(path))
    assert t2.cache == tuner.cache
    wfs.engine.nchunk_user = None
    assert t2.tune(wfs, cs.amp, cs.scale, torch.zeros((1, 32, 32), device='cuda')) == best
    assert wfs.engine.nchunk_user == best['nchunk']
