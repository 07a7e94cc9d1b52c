"""GPU: the drop-in object chain (src**tel*dm*wfs) against the oracle at config 1.

Reference: Telescope.relay (Telescope.py:199-224), DeformableMirror.coefs /
relay (DeformableMirror.py:427-457, 651-686), Pyramid.__init__ / wfs_measure /
signalProcessing (Pyramid.py:213-366, 541-610, 666-854).
Tolerances: float32 GPU path vs float64 oracle, rel-L2 <= 1e-5 on slopes;
<= 1e-6 on the DM OPD.
"""
import numpy as np
import pytest

from tests.helpers import rel_l2, config1_system, random_phase

pytestmark = pytest.mark.gpu


def _config1_objects():
    from oopao_amd import Telescope, Source, DeformableMirror, Pyramid
    tel = Telescope(80, 8, 1e-3, centralObstruction=0)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    dm = DeformableMirror(tel, 20, 0.35, print_dm_properties=False)
    wfs = Pyramid(20, tel, 3, 0.5, n_pix_separation=20, n_pix_edge=10, nTheta_user_defined=48)
    return tel, ngs, dm, wfs


def test_dm_opd_matches_oracle():
    from oracle.dm import DMOracle
    tel, ngs, dm, wfs = _config1_objects()
    o = DMOracle(80, 8, 20, 0.35)
    assert dm.nValidAct == o.nValidAct == 357
    coefs = np.random.RandomState(2).normal(0, 20e-9, 357)
    dm.coefs = coefs
    ref = o.opd(o.modes(), coefs)
    err = rel_l2(dm.OPD, ref)
    print('separable OPD rel_l2', err)
    assert err < 1e-6
    dm.use_dense = True
    dm.coefs = coefs
    err = rel_l2(dm.OPD, ref)
    print('dense OPD rel_l2', err)
    assert err < 1e-6
    # batched coefficients -> cube [n, n, B]
    C = np.random.RandomState(3).normal(0, 20e-9, (357, 5))
    dm.coefs = C
    assert dm.OPD.shape == (80, 80, 5)
    assert rel_l2(dm.OPD, o.opd(o.modes(), C)) < 1e-6


def test_pyramid_init_matches_oracle():
    tel, ngs, dm, wfs = _config1_objects()
    o = config1_system()
    assert wfs.resolution == o.geom.N == 320
    assert wfs.cam.resolution == 80
    assert wfs.nSignal == o.nSignal
    assert np.array_equal(wfs.validI4Q, o.validI4Q)
    assert rel_l2(wfs.referenceSignal_2D, o.ref2d) < 1e-5
    # the final flat measurement in the constructor gives zero slopes
    assert np.max(np.abs(wfs.signal)) < 1e-5


def test_chain_signal_matches_oracle():
    from oracle.dm import DMOracle
    tel, ngs, dm, wfs = _config1_objects()
    o = config1_system()
    od = DMOracle(80, 8, 20, 0.35)
    coefs = np.random.RandomState(2).normal(0, 20e-9, 357)
    dm.coefs = coefs
    ngs ** tel * dm * wfs
    opd = od.opd(od.modes(), coefs)
    s_ref = o.measure_opd(opd)[1]
    err = rel_l2(wfs.signal, s_ref)
    print('chain slopes rel_l2 vs fp64 oracle', err)
    assert err <= 1e-5
    # batched (3-D) phase: three screens through the same chain
    C = np.random.RandomState(4).normal(0, 50e-9, (357, 3))
    dm.coefs = C
    ngs ** tel * dm * wfs
    assert wfs.signal.shape == (wfs.nSignal, 3)
    for j in range(3):
        s_ref = o.measure_opd(od.opd(od.modes(), C[:, j]))[1]
        assert rel_l2(wfs.signal[:, j], s_ref) <= 1e-5


def test_focal_plane_camera_matches_oracle():
    # wfs * wfs.focal_plane_camera (Pyramid.py:1178-1203): resolution (3*4+12)*zpf = 96 px
    from oracle.dm import DMOracle
    tel, ngs, dm, wfs = _config1_objects()
    o = config1_system()
    od = DMOracle(80, 8, 20, 0.35)
    coefs = np.random.RandomState(5).normal(0, 30e-9, 357)
    dm.coefs = coefs
    ngs ** tel * dm * wfs
    wfs * wfs.focal_plane_camera
    fr = wfs.focal_plane_camera.frame
    assert fr.shape == (96, 96)
    opd = od.opd(od.modes(), coefs)
    ref = o.focal(opd * o.pupil * 2 * np.pi / o.wavelength, 96)
    err = rel_l2(fr, ref)
    print('focal camera rel_l2 vs fp64 oracle', err)
    assert err <= 1e-5


def test_get_modulation_frame_crop_and_norm():
    # Pyramid.get_modulation_frame (Pyramid.py:928-944) on a full N x N focal-camera frame: crop of
    # +-radius lambda/D (radius * zeroPaddingFactor pixels) about the centre, normalised to its maximum
    tel, ngs, dm, wfs = _config1_objects()
    o = config1_system()
    N = wfs.resolution
    wfs.focal_plane_camera.resolution = N
    ngs ** tel * dm * wfs
    wfs * wfs.focal_plane_camera
    ref = o.focal(np.zeros((80, 80)), N)
    for radius in (6, 10):
        zoom = wfs.get_modulation_frame(radius=radius)
        t = int(N / 2 - radius * wfs.zeroPaddingFactor)
        r = ref[t:-t, t:-t] / ref[t:-t, t:-t].max()
        assert zoom.shape == (N - 2 * t, N - 2 * t)
        assert zoom.max() == 1.0
        assert rel_l2(zoom, r) <= 1e-5
    # radius beyond the field of view: the whole frame, unnormalised on request
    full = wfs.get_modulation_frame(radius=N, norma=False)
    assert full.shape == (N, N)
    assert rel_l2(full, wfs.focal_plane_camera.frame) == 0.0


@pytest.mark.parametrize('mode', ['slopesMaps_incidence_flux', 'slopesMaps_camera_flux', 'fullFrame_camera_flux',
                                  'fullFrame_incidence_flux', 'fullFrame_sum_flux', 'fullFrame'])
def test_post_processing_modes(mode):
    # signalProcessing's other modes (Pyramid.py:856-926) on the config-1 sensor: the signal of the
    # calibration (flat) wavefront is zero, the modes agree with the reference's expressions evaluated on
    # the camera frame, and the slopes modes share the raw quadrant differences of 'slopesMaps'
    from oopao_amd import Pyramid
    tel, ngs, dm, wfs0 = _config1_objects()
    wfs = Pyramid(20, tel, 3, 0.5, n_pix_separation=20, n_pix_edge=10, nTheta_user_defined=48, postProcessing=mode)
    tel.resetOPD()
    ngs * tel * wfs
    s_flat = np.asarray(wfs.signal)
    assert np.abs(s_flat).max() <= 1e-6 * max(1.0, np.abs(wfs.referenceSignal_2D).max())
    opd = np.random.RandomState(9).randn(80, 80) * 40e-9
    tel.OPD_no_pupil = opd
    ngs * tel * wfs
    frame = wfs.cam.frame
    ref2d = wfs.referenceSignal_2D
    valid = np.asarray(wfs.validSignal).astype(bool)
    sub_flux = ngs.nPhoton * tel.samplingTime * (tel.D / wfs.nSubap) ** 2
    if mode.startswith('slopesMaps'):
        I = [wfs.grabQuadrant(k, frame) * wfs.validI4Q for k in (1, 2, 3, 4)]
        raw = np.concatenate((I[0] - I[1] + I[3] - I[2], I[0] - I[3] + I[1] - I[2]))
        norma = sub_flux if mode == 'slopesMaps_incidence_flux' else frame.mean()
        exp2d = (raw / norma - ref2d) * wfs.slopesUnits
        # the same raw differences as the default mode, up to the normalisation
        tel.OPD_no_pupil = opd
        ngs * tel * wfs0
        raw0 = (wfs0.signal_2D / wfs0.slopesUnits + wfs0.referenceSignal_2D) * wfs0.norma
        I4Q = sum(I)
        assert abs(wfs0.norma / I4Q[wfs.validI4Q].mean() - 1) < 1e-6
        assert rel_l2(raw0, raw) <= 1e-5
    else:
        norma = {'fullFrame_camera_flux': frame.mean(), 'fullFrame_incidence_flux': sub_flux / 4,
                 'fullFrame_sum_flux': frame.sum(), 'fullFrame': frame[valid].sum()}[mode]
        exp2d = frame / norma - ref2d
    assert rel_l2(wfs.signal_2D, exp2d) <= 1e-5
    assert rel_l2(wfs.signal, exp2d[valid]) <= 1e-5
    assert abs(wfs.norma / norma - 1) < 1e-6


@pytest.mark.parametrize('nTheta_user', [48, None])
def test_modulation_and_light_ratio_setters(nTheta_user):
    # modulation setter (Pyramid.py:1107-1154: new path, recalibrated reference) and lightRatio setter
    # (re-threshold of the stored I4Q) against the oracle built with the new values
    from oopao_amd import Pyramid
    from oracle.pwfs import PyramidOracle, telescope_pupil, flux_map, n_photon
    tel, ngs, dm, _ = _config1_objects()
    wfs = Pyramid(20, tel, 3, 0.5, n_pix_separation=20, n_pix_edge=10, nTheta_user_defined=nTheta_user)
    pup = telescope_pupil(80)
    flux = flux_map(pup, n_photon('I3', 8), 1e-3, 8.0)
    opd = np.random.RandomState(11).randn(80, 80) * 30e-9
    o = None
    for mod, lr in ((5, 0.5), (5, 0.2), (2, 0.2)):
        if mod != wfs.modulation:
            wfs.modulation = mod
            o = PyramidOracle(80, 20, mod, lr, pup, flux, 0.8e-6, n_pix_separation=20, n_pix_edge=10,
                              nTheta_user_defined=nTheta_user)
        if lr != wfs.lightRatio:
            # the setter re-thresholds the stored I4Q and keeps the reference slopes of the last calibration
            wfs.lightRatio = lr
            o.validI4Q = o.I4Q >= lr * o.I4Q.max()
            o.validSignal = np.concatenate((o.validI4Q, o.validI4Q))
            o.nSignal = int(o.validSignal.sum())
        assert wfs.nTheta == len(o.path)
        assert np.array_equal(np.asarray(wfs.validI4Q, bool), o.validI4Q)
        assert wfs.nSignal == o.nSignal
        tel.OPD_no_pupil = opd
        ngs * tel * wfs
        err = rel_l2(wfs.signal, o.measure_opd(opd)[1])
        print('modulation', mod, 'lightRatio', lr, 'nTheta', wfs.nTheta, 'rel_l2', err)
        assert err <= 1e-5


def test_error_paths_raise_oopao_error():
    # DeformableMirror.py:683-685 exits the interpreter on a wrong coefficient count: the drop-in raises;
    # Pyramid.py:667-670 raises on a source whose wavelength differs from the calibration wavelength;
    # a phase-screen interaction matrix of the wrong map size is rejected before any device work
    from oopao_amd import Source, InteractionMatrixFromPhaseScreen
    from oopao_amd.tools import OopaoError
    tel, ngs, dm, wfs = _config1_objects()
    with pytest.raises(OopaoError):
        dm.coefs = np.zeros(356)
    with pytest.raises(OopaoError):
        dm.coefs = 1.0
    dm.coefs = 0
    ngs ** tel * dm * wfs
    s0 = wfs.signal.copy()
    other = Source('K', 8, display_properties=False)
    with pytest.raises(OopaoError):
        other ** tel * dm * wfs
    with pytest.raises(OopaoError):
        InteractionMatrixFromPhaseScreen(ngs, None, tel, wfs, np.zeros((40, 40, 2)), 1e-8, print_time=False)
    # the objects stay usable after the errors
    ngs ** tel * dm * wfs
    assert np.array_equal(wfs.signal, s0)


def test_workspace_failure_halves_the_batch(monkeypatch):
    # a workspace that cannot be allocated (forced from the host by the byte cap) makes the Pyramid
    # measure fewer screens per call instead of failing; the frames equal the unconstrained ones
    import torch
    tel, ngs, dm, wfs = _config1_objects()
    C = np.random.RandomState(6).normal(0, 40e-9, (357, 8))
    dm.coefs = C
    ngs ** tel * dm * wfs
    ref = wfs.cam._frame.clone() if hasattr(wfs.cam, '_frame') else None
    s_ref = wfs.signal.copy()
    tel2, ngs2, dm2, wfs2 = _config1_objects()
    per = wfs2.engine.bytes_per_screen()
    monkeypatch.setenv('OOPAO_AMD_MAX_WORKSPACE_BYTES', str(int(2.5 * per)))
    dm2.coefs = C
    ngs2 ** tel2 * dm2 * wfs2
    monkeypatch.delenv('OOPAO_AMD_MAX_WORKSPACE_BYTES')
    assert wfs2.max_batch_bytes <= 2 * per
    assert wfs2.signal.shape == (wfs2.nSignal, 8)
    assert rel_l2(wfs2.signal, s_ref) <= 1e-6
    if ref is not None:
        assert torch.allclose(wfs2.cam._frame, ref, rtol=1e-5, atol=1e-6 * float(ref.abs().max()))
    # a single screen that cannot get its workspace is an error, not a silent result
    # (the Pyramid's own calibration measures one screen)
    monkeypatch.setenv('OOPAO_AMD_MAX_WORKSPACE_BYTES', str(per // 2))
    from oopao_amd.tools import WorkspaceAllocationError
    with pytest.raises(WorkspaceAllocationError):
        _config1_objects()


def test_config1_exact_input():
    # config 1 as specified: the seed-1 von Karman screen of a one-layer atmosphere (r0 0.15 m, L0 25 m)
    # plus the OPD of the 21 x 21 DM driven by RandomState(2).normal(0, 20e-9, 357), through
    # ngs**atm*tel*dm*wfs. The screen is ft_sh_phase_screen on the layer's 84^2 support (8.4 m, 0.1 m
    # pixels, seed 1) cropped to the pupil footprint [2:82, 2:82] and scaled by 500e-9 / 2 pi
    # (Atmosphere.py:779-810, 589-622; oracle/atmosphere restates the generator)
    from oopao_amd import Atmosphere
    from oracle import atmosphere as oat
    from oracle.dm import DMOracle
    tel, ngs, dm, wfs = _config1_objects()
    atm = Atmosphere(tel, 0.15, 25, windSpeed=[10], fractionalR0=[1], windDirection=[0], altitude=[0])
    atm.initializeAtmosphere(tel)
    atm.generateNewPhaseScreen(1)
    screen = oat.ft_sh_phase_screen(0.15, 25, 84, 8.4 / 84, seed=1)[2:82, 2:82] * 500e-9 / (2 * np.pi)
    assert rel_l2(atm.OPD, screen) < 1e-6
    coefs = np.random.RandomState(2).normal(0, 20e-9, 357)
    dm.coefs = coefs
    ngs ** atm * tel * dm * wfs
    od = DMOracle(80, 8, 20, 0.35)
    s_ref = config1_system().measure_opd(screen + od.opd(od.modes(), coefs))[1]
    err = rel_l2(wfs.signal, s_ref)
    print('config-1 exact input: slopes rel-L2 %.2e vs float64 oracle (nSignal %d)' % (err, wfs.nSignal))
    assert err <= 1e-5
