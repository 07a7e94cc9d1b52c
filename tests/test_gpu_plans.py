"""GPU: every compiled FFT plan of the separable PWFS gives the oracle's frames (Pyramid.py:612-757).

At the sensor sizes a batched calibration mixes (n = 80, 160, 256, 320 -> N = 320, 640, 1024, 1280) the
library compiles a second plan for the rows and / or columns pass (another lane count per transform, hence
another radix plan and occupancy). Each plan, alone and combined with the other pass's plans, is checked
against the float64 oracle frame at rel-L2 <= 1e-5 (the float32 bar of the default plan), whole and
slopes-only (column bands), and against the default plan's frame at 2e-6 (float rounding of a different
radix order). A plan that is not compiled is refused. The autotuner records the plans it picked.
"""
import numpy as np
import pytest

from tests.helpers import rel_l2, random_phase

pytestmark = pytest.mark.gpu


def _system(n, nsub, T):
    from oracle import pwfs as opw
    pup = opw.telescope_pupil(n)
    flux = opw.flux_map(pup, opw.n_photon('I3', 8), 1e-3, 8.0)
    return opw.PyramidOracle(n, nsub, 3, 0.5, pup, flux, 0.8e-6, n_pix_separation=nsub, n_pix_edge=nsub // 2,
                             nTheta_user_defined=T, n_jobs=8)


def _engine(o):
    import torch
    from oopao_amd.pwfs_engine import PWFSEngine
    from oracle.pwfs import tip_tilt
    g = o.geom
    dev = torch.device('cuda', 0)
    eng = PWFSEngine(g.n, g.N, g.cam, dev)
    tip, tilt = tip_tilt(g.n)
    eng.set_modulation(o.path, tip[0, :], tilt[:, 0])
    eng.set_mask(o.m)
    amp = torch.from_numpy(np.sqrt(o.flux / o.nTheta).astype(np.float32)).to(dev)
    return eng, amp


# (n, subapertures, modulation points): the config-3 sensors' sizes, with few points for the oracle
@pytest.mark.parametrize('n,nsub,T', [(80, 20, 12), (160, 40, 8), (256, 64, 6), (320, 80, 4)])
def test_every_plan_matches_oracle(n, nsub, T):
    import torch
    from oopao_amd.tools import OopaoError
    o = _system(n, nsub, T)
    eng, amp = _engine(o)
    rows, cols = eng.plans()
    print('n=%d N=%d plans: rows %s, columns %s' % (n, eng.N, rows, cols), flush=True)
    assert len(rows) + len(cols) >= 3      # a second plan for at least one pass at these sizes
    ph = random_phase(n, 11, 1.0, o.pupil)
    ref = o.frame(ph)
    x = torch.from_numpy(ph[None].astype(np.float32)).to(amp.device).contiguous()
    c, e, w = eng.cam // 2, nsub // 2, nsub
    bands = [(c - e - w, c - e), (c + e, c + e + w)]
    keep = np.zeros(eng.cam, bool)
    for a, b in bands:
        keep[a:b] = True
    base = None
    for lr in rows:
        for lc in cols:
            eng.set_plans(lr, lc)
            got = eng.frames(x, 1.0, amp).cpu().numpy()[0]
            band = eng.frames(x, 1.0, amp, bands=bands).cpu().numpy()[0]
            err = rel_l2(got, ref)
            print('  rows %d lanes, columns %d lanes: rel_l2 vs fp64 oracle %.2e' % (lr, lc, err), flush=True)
            assert err <= 1e-5
            assert np.array_equal(band[:, keep], got[:, keep]) and not band[:, ~keep].any()
            if base is None:
                base = got
            else:
                assert rel_l2(got, base) <= 2e-6
    eng.set_plans(0, 0)
    with pytest.raises(OopaoError):
        eng.set_plans(8, 0)
    torch.cuda.synchronize()


def test_autotuner_records_plans():
    import torch
    from oopao_amd import Telescope, Source, Pyramid
    from oopao_amd.calibration.batched import PWFSAutotuner
    tel = Telescope(160, 8.0, 1e-3)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    wfs = Pyramid(40, tel, 3, 0.5, n_pix_separation=40, n_pix_edge=20, nTheta_user_defined=48)
    tuner = PWFSAutotuner(batch_candidates=(4, 8))
    opd = torch.zeros((1, 160, 160), dtype=torch.float32, device=wfs.device)
    best = tuner.tune(wfs, wfs._amp(), 2 * np.pi / ngs.wavelength, opd)
    print('autotuner choice at n=160:', best, flush=True)
    rows, cols = wfs.engine.plans()
    assert best['lanes_rows'] in rows and best['lanes_cols'] in cols
    assert set(best['plan_ms']) == {'rows_%d' % v for v in rows} | {'cols_%d' % v for v in cols}
    assert (wfs.engine.lanes_rows, wfs.engine.lanes_cols) == (best['lanes_rows'], best['lanes_cols'])
