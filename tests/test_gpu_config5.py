"""GPU: config 5 at its workload -- the ELT-scale pupil (39 m, 30 % central obstruction, 480 px).

* PWFS, 120 x 120 subapertures at 4 px (N = 1920, 96 modulation points; the survey's 100 x 100 at 4.8 px
  raises in the reference, Pyramid.py:216-217): the whole frame and the slopes of a seeded 60 nm RMS
  screen against the float64 oracle with all 96 points (oracle/pwfs.PyramidOracle, 16 threads),
  rel-L2 <= 1e-5; the frame's total equals the pupil flux (Parseval, |mask| = 1).
* The 80 x 80 DM's influence functions as a BF16 matrix [230400 x nAct] on the matrix-core fragment
  layout applied to one command (GEMV, BF16 and split BF16), against float64 rows from the Gaussian
  formula (oracle/dm.DMOracle.modes_rows) on 512 sampled pupil pixels. Bounds per output row r with
  S_r = sum_k |A_rk x_k|: BF16 |err| <= (2^-8 + K 2^-24) S_r + K 2^-126 max|x| (the matrix rounded to 8
  bits, float32 accumulation, and entries below float32's smallest normal number -- Gaussian tails far from
  every actuator -- held as zero); the unit roundoff of BF16's 8 significant bits is 2^-8, and split BF16
  (hi + lo) keeps the matrix to 2^-16.
* A BF16 reconstructor [nAct x 15700] applied to one slope vector, against the float64 product of its
  float32 matrix on 512 sampled rows, same bounds.
"""
import numpy as np
import pytest
import torch

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def _elt():
    from oopao_amd import Telescope, Source
    tel = Telescope(480, 39.0, 1e-3, centralObstruction=0.3)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    return tel, ngs


def test_elt_pwfs_full_frame_vs_oracle():
    from oopao_amd import Pyramid
    from oracle import pwfs as opw
    from tests.helpers import random_phase
    print('ELT PWFS: n=480, N=1920, 96 points', flush=True)
    tel, ngs = _elt()
    wfs = Pyramid(120, tel, 3, 0.5, n_pix_separation=120, n_pix_edge=60, nTheta_user_defined=96)
    assert wfs.engine.N == 1920 and wfs.engine.mode == 'fast'
    flux = opw.flux_map(tel.pupil, ngs.nPhoton, tel.samplingTime, tel.D)
    o = opw.PyramidOracle(480, 120, 3, 0.5, tel.pupil, flux, ngs.wavelength, n_pix_separation=120, n_pix_edge=60,
                          nTheta_user_defined=96, n_jobs=16)
    try:
        assert np.array_equal(wfs.validI4Q, o.validI4Q)
        opd = random_phase(480, 21, 1.0, tel.pupil) * 60e-9
        od = torch.from_numpy((opd * tel.pupil).astype(np.float32)).to(wfs.device)[None]
        frame = wfs.frames_dev(od)[0].double().cpu().numpy()
        assert abs(frame.sum() / flux.sum() - 1) < 1e-5
        ref = o.frame(opd * tel.pupil * 2 * np.pi / o.wavelength)
        e_frame = rel_l2(frame, ref)
        s2d, s = wfs.slopes_engine.slopes(torch.from_numpy(frame.astype(np.float32)).to(wfs.device)[None])
        s_ref = o.measure_opd(opd)[1]
        e_s = rel_l2(s[0].double().cpu().numpy(), s_ref)
        print('ELT frame rel-L2 %.2e, slopes rel-L2 %.2e vs float64 oracle (all 96 points)' % (e_frame, e_s))
        assert e_frame <= 1e-5 and e_s <= 1e-5
    finally:
        if o._pool is not None:
            o._pool.shutdown()


def _row_bound_check(y, exact, S, K, lo_bits, xmax):
    err = np.abs(y - exact)
    bound = (2.0 ** -lo_bits + K * 2.0 ** -24) * S + K * 2.0 ** -126 * xmax
    return float(np.max(err / bound)), rel_l2(y, exact)


def test_elt_bf16_contractions_vs_float64():
    from oopao_amd import DeformableMirror
    from oopao_amd.contract import Contraction, PackedBF16
    from oracle.dm import DMOracle
    tel, ngs = _elt()
    dm = DeformableMirror(tel, 80, 0.35, print_dm_properties=False)
    modes = dm._build_modes()
    R, K = modes.shape
    od = DMOracle(480, 39.0, 80, 0.35, central_obstruction=0.3)
    assert (R, K) == (480 * 480, od.nValidAct)
    print('ELT DM: influence matrix [%d x %d]' % (R, K), flush=True)
    dev = modes.device
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    x = (torch.randn(K, device=dev, generator=g) * 20e-9).contiguous()
    rows = np.random.RandomState(8).choice(R, 512, replace=False)
    A64 = od.modes_rows(rows)
    xd = x.double().cpu().numpy()
    exact = A64 @ xd
    S = np.abs(A64) @ np.abs(xd)
    C = Contraction(PackedBF16(modes, split=True))
    rows_t = torch.from_numpy(rows).to(dev)
    for prec, bits in (('bf16', 8), ('bf16x3', 16)):
        y = C.gemv(x, precision=prec)
        ratio, e = _row_bound_check(y[rows_t].double().cpu().numpy(), exact, S, K, bits, np.abs(xd).max())
        print('IF GEMV %s: rel-L2 %.2e vs float64 rows, worst error / bound %.3f' % (prec, e, ratio))
        assert ratio <= 1.0
    del C, modes
    torch.cuda.empty_cache()
    nsig = 15700
    Rm = torch.randn(K, nsig, device=dev, generator=g) * 1e-3
    s = torch.randn(nsig, device=dev, generator=g)
    rrows = torch.from_numpy(np.random.RandomState(9).choice(K, 512, replace=False)).to(dev)
    Rs = Rm[rrows].double()
    exact = (Rs @ s.double()).cpu().numpy()
    S = (Rs.abs() @ s.double().abs()).cpu().numpy()
    C = Contraction(PackedBF16(Rm, split=True))
    for prec, bits in (('bf16', 8), ('bf16x3', 16)):
        c = C.gemv(s, precision=prec)
        ratio, e = _row_bound_check(c[rrows].double().cpu().numpy(), exact, S, nsig, bits, float(s.abs().max()))
        print('reconstructor MVM %s: rel-L2 %.2e vs float64, worst error / bound %.3f' % (prec, e, ratio))
        assert ratio <= 1.0


def test_elt_pwfs_4p8_px_per_subaperture_vs_oracle():
    # the configuration as specified: 100 x 100 subapertures on the 480 px pupil (4.8 px each), N = 1920,
    # 96 points; the reference refuses this sampling (Pyramid.py:216-217), so the drop-in takes it with
    # fractional_sampling=True: the separable transforms give the unbinned frame, integrated onto the
    # 400 x 400 detector by area; the oracle is the float64 restatement with the same area integration
    # (oracle/pwfs.bin_area, pinned by exact sub-pixel refinement)
    from oopao_amd import Pyramid
    from oracle import pwfs as opw
    from tests.helpers import random_phase
    print('ELT PWFS at 4.8 px per subaperture: n=480, N=1920, cam=400, 96 points', flush=True)
    tel, ngs = _elt()
    wfs = Pyramid(100, tel, 3, 0.5, n_pix_separation=100, n_pix_edge=50, nTheta_user_defined=96,
                  fractional_sampling=True)
    assert wfs.engine.N == 1920 and wfs.cam.resolution == 400 and wfs.engine.fractional
    assert wfs.slope_bands() is None
    flux = opw.flux_map(tel.pupil, ngs.nPhoton, tel.samplingTime, tel.D)
    o = opw.PyramidOracle(480, 100, 3, 0.5, tel.pupil, flux, ngs.wavelength, n_pix_separation=100, n_pix_edge=50,
                          nTheta_user_defined=96, n_jobs=16, fractional=True)
    try:
        assert np.array_equal(wfs.validI4Q, o.validI4Q) and wfs.nSignal == o.nSignal
        opd = random_phase(480, 22, 1.0, tel.pupil) * 60e-9
        od = torch.from_numpy((opd * tel.pupil).astype(np.float32)).to(wfs.device)[None]
        frame = wfs.frames_dev(od)[0].double().cpu().numpy()
        assert abs(frame.sum() / flux.sum() - 1) < 1e-5
        raw_ref = o.raw(opd * tel.pupil * 2 * np.pi / o.wavelength)
        ref = opw.bin_area(raw_ref, 400)
        e_frame = rel_l2(frame, ref)
        raw = wfs.engine.raw(od, 2 * np.pi / ngs.wavelength, wfs._amp())[0].double().cpu().numpy()
        e_raw = rel_l2(raw, raw_ref)
        s2d, s = wfs.slopes_engine.slopes(torch.from_numpy(frame.astype(np.float32)).to(wfs.device)[None])
        e_s = rel_l2(s[0].double().cpu().numpy(), o.measure_opd(opd)[1])
        print('4.8 px: raw rel-L2 %.2e, frame %.2e, slopes %.2e vs float64 oracle' % (e_raw, e_frame, e_s))
        assert e_raw <= 1e-5 and e_frame <= 1e-5 and e_s <= 1e-5
    finally:
        if o._pool is not None:
            o._pool.shutdown()


@pytest.mark.parametrize('N,cam', [(1920, 400), (1280, 320), (320, 80), (264, 44)])
def test_bin_area_kernel(N, cam):
    import ctypes
    from oopao_amd import _lib
    from oracle import pwfs as opw
    dev = torch.device('cuda', 0)
    rs = np.random.RandomState(N)
    raw = rs.rand(2, N, N).astype(np.float32)
    out = torch.empty((2, cam, cam), dtype=torch.float32, device=dev)
    r = torch.from_numpy(raw).to(dev)
    _lib.call('oopao_bin_area', ctypes.c_void_p(r.data_ptr()), 2, N, cam, ctypes.c_void_p(out.data_ptr()),
              ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
    got = out.double().cpu().numpy()
    for b in range(2):
        assert rel_l2(got[b], opw.bin_area(raw[b].astype(np.float64), cam)) <= 1e-7
