"""GPU: slopes-only frames (the columns the slopesMaps signal reads, Pyramid.py:838-853, 946-961).

A frame computed for the quadrant column bands must equal the whole frame bit for bit in those columns
and be zero elsewhere, so the slopes are bit for bit those of the whole frame; at config 2 (full size,
72 points) they also match the float64 oracle within 1e-5. Through the drop-in Pyramid, a measurement
on a noiseless camera takes this path and cam.frame computes the whole frame when read.
"""
import ctypes

import numpy as np
import pytest
import torch

from tests.helpers import rel_l2, random_phase

pytestmark = pytest.mark.gpu


def _pyramid(n, nSubap, T):
    from oopao_amd import Telescope, Source, Pyramid
    tel = Telescope(n, 8, 1e-3)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    wfs = Pyramid(nSubap, tel, 3, 0.5, n_pix_separation=nSubap, n_pix_edge=nSubap // 2, nTheta_user_defined=T)
    return tel, ngs, wfs


@pytest.mark.parametrize('n,nSubap,T', [(80, 20, 48), (320, 80, 72)])
def test_band_frames_equal_whole_frames(n, nSubap, T):
    tel, ngs, wfs = _pyramid(n, nSubap, T)
    eng, se = wfs.engine, wfs.slopes_engine
    bands = wfs.slope_bands()
    quad, w = wfs._quadrant_origins()
    assert bands == [(quad[4], quad[4] + w), (quad[5], quad[5] + w)]
    print('n=%d: detector columns %s of %d' % (n, bands, eng.cam), flush=True)
    B = 3
    ph = np.stack([random_phase(n, 40 + k, 1.5, tel.pupil) for k in range(B)]) * ngs.wavelength / (2 * np.pi)
    opd = torch.from_numpy(ph.astype(np.float32)).to(wfs.device).contiguous()
    amp = wfs._amp()
    scale = 2 * np.pi / ngs.wavelength
    full = eng.frames(opd, scale, amp)
    part = eng.frames(opd, scale, amp, bands=bands)
    cols = np.zeros(eng.cam, bool)
    for a, b in bands:
        cols[a:b] = True
    cols_t = torch.from_numpy(cols).to(wfs.device)
    assert torch.equal(part[:, :, cols_t], full[:, :, cols_t])
    assert torch.count_nonzero(part[:, :, ~cols_t]) == 0
    s2d_f, s_f = se.slopes(full)
    s2d_p, s_p = se.slopes(part)
    assert torch.equal(s2d_f, s2d_p) and torch.equal(s_f, s_p)
    # the split passes with the same bands
    buf = eng.buffers(B)
    out = torch.empty_like(full)
    eng.rows(opd, scale, amp, buf, bands)
    eng.cols(buf, out, bands)
    assert torch.equal(out, part)
    if n == 320:
        # config 2 at full size against the float64 oracle
        from tests.helpers import config2_system
        o = config2_system(n_jobs=16)
        try:
            assert np.array_equal(wfs.validI4Q, o.validI4Q)
            ref = o.measure_opd(ph[0].astype(np.float64))[1]
            # the drop-in's reference slopes (float32 flat frame) against the oracle's (float64)
            got = s_p[0].double().cpu().numpy()
            err = rel_l2(got, ref)
            print('config 2 slopes-only vs float64 oracle: rel-L2 %.3e' % err)
            assert err <= 1e-5
        finally:
            if o._pool is not None:
                o._pool.shutdown()


def test_dropin_measurement_is_slopes_only_and_frame_is_whole_on_read():
    from oopao_amd import DeformableMirror
    tel, ngs, wfs = _pyramid(80, 20, 48)
    dm = DeformableMirror(tel, 20, 0.35, print_dm_properties=False)
    dm.coefs = np.random.RandomState(2).normal(0, 20e-9, dm.nValidAct)
    ngs ** tel * dm * wfs
    assert wfs.cam._lazy is not None                  # the measurement computed the quadrant columns only
    pruned = wfs.cam._frame.clone()
    sig = wfs.signal.copy()
    # the whole frame of the same OPD, and its slopes
    opd, _, amp = wfs._last_opd
    full = wfs.engine.frames(opd, 2 * np.pi / ngs.wavelength, amp)[0]
    _, s_full = wfs.slopes_engine.slopes(full[None])
    assert np.array_equal(sig, s_full[0].double().cpu().numpy())
    # reading cam.frame computes the whole frame once
    fr = wfs.cam.frame
    assert wfs.cam._lazy is None
    assert np.array_equal(fr, full.double().cpu().numpy())
    assert torch.count_nonzero(pruned) < torch.count_nonzero(full)
    assert np.array_equal(wfs.frames, fr)
    # a noisy camera needs the whole frame: no slopes-only measurement
    wfs.cam.photonNoise = True
    wfs.cam.seed_noise(5)
    ngs ** tel * dm * wfs
    assert wfs.cam._lazy is None
    wfs.cam.photonNoise = False


def test_band_and_buffer_argument_checks():
    from oopao_amd import _lib
    from oopao_amd.tools import OopaoError
    from oopao_amd.pwfs_engine import column_bands
    tel, ngs, wfs = _pyramid(80, 20, 48)
    eng = wfs.engine
    assert list(column_bands([(50, 70), (10, 30)], 80)) == [10, 30, 50, 70]
    assert list(column_bands([(10, 30), (20, 40)], 80)) == [10, 40, 80, 80]
    with pytest.raises(OopaoError):
        column_bands([(0, 10), (20, 30), (40, 50)], 80)
    with pytest.raises(OopaoError):
        column_bands([(70, 90)], 80)
    # the library rejects out-of-range bands and undersized buffers before any device work
    L = _lib.lib()
    dummy = ctypes.c_void_p(256)
    bad = (ctypes.c_int * 4)(30, 10, 50, 70)
    rc = L.oopao_pwfs_cols(dummy, 1 << 30, dummy, 80, 320, 48, 1, 8, 0, bad, dummy, 1 << 30, dummy, None)
    assert rc == 1 and b'bands' in L.oopao_last_error()
    rc = L.oopao_pwfs_cols(dummy, 100, dummy, 80, 320, 48, 1, 8, 0, None, dummy, 1 << 30, dummy, None)
    assert rc == 4
    # a buffers dict sized for another number of modulation points is refused
    buf = eng.buffers(1)
    buf['T'] = eng.T + 1
    with pytest.raises(OopaoError):
        eng.cols(buf, torch.empty((1, eng.cam, eng.cam), device=wfs.device))


@pytest.mark.parametrize('n,nSubap,T', [(80, 20, 48), (320, 80, 72)])
def test_rows_pass_work_split_does_not_change_frames(n, nSubap, T):
    """The rows pass deals its (screen, tip table, row block) items to the resident workgroups with a
    stride, in as many rounds as the cap of tables per workgroup needs: every cap (one table per
    workgroup and many rounds, up to all tables in one round), slopes-only or whole, gives the same
    frames bit for bit (the intermediate does not depend on the split)."""
    tel, ngs, wfs = _pyramid(n, nSubap, T)
    eng = wfs.engine
    bands = wfs.slope_bands()
    B = 3
    ph = np.stack([random_phase(n, 60 + k, 1.5, tel.pupil) for k in range(B)]) * ngs.wavelength / (2 * np.pi)
    opd = torch.from_numpy(ph.astype(np.float32)).to(wfs.device).contiguous()
    amp = wfs._amp()
    scale = 2 * np.pi / ngs.wavelength
    for bd in (bands, None):
        ref = None
        for ppt in (1, 3, eng.P):
            eng.pts_per_task_user = ppt
            f = eng.frames(opd, scale, amp, bands=bd).clone()
            if ref is None:
                ref = f
            print('n=%d bands=%s ppt=%d: max |diff| %.3g' % (n, bd is not None, ppt, float((f - ref).abs().max())),
                  flush=True)
            assert torch.equal(f, ref)
    eng.pts_per_task_user = None
