"""Benchmark: modulated Pyramid WFS frames/s at 80x80 subapertures, 72 modulation points.

Workload (one "step" = one PWFS frame): 8 m pupil sampled on 320 x 320 px, I3
(800 nm) magnitude-8 source, Pyramid(nSubap=80, modulation=3 lambda/D,
n_pix_separation=80, n_pix_edge=40, nTheta_user_defined=72) -> 1280^2 padded
fields, 320^2 detector, slopesMaps signal. Input: a seeded synthetic OPD map
(smooth random screen, fixed seed) resident in HBM before the timed region;
output: the slope vector. The timed step runs the rows pass, the columns pass
(+ chunk reduction) and the slopes kernel on one stream, for the detector columns
the slopesMaps signal reads (the quadrant bands, half of the 320 columns); the
whole 320x320 frame (the same kernels over every column) is timed beside it.

Prints one JSON line. ``--gpus N`` with N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# config-2 geometry
N_RES, N_SUBAP, SEP, EDGE, MOD, T_POINTS, DIAM, TS = 320, 80, 80, 40, 3, 72, 8.0, 1e-3
FP32_PEAK_TFLOPS = 157.3   # MI355X published FP32 peak (packed FP32 issue)
VALU_PEAK_TFLOPS = 78.6    # FP32 vector FMA without packed instructions (256 CU x 64 lanes x 2 x 2.4 GHz); the
                           # build disables packed FP32 (-packed-fp32-ops), so this is the issue ceiling
HBM_PEAK_GBS = 8000.0
BF16_MFMA_PEAK_TFLOPS = 2500.0   # dense BF16 matrix-core peak


def synthetic_opd(n, seed=10, rms_nm=150.0):
    rs = np.random.RandomState(seed)
    w = rs.randn(n, n)
    k = np.fft.fftfreq(n)
    f = np.sqrt(k[:, None] ** 2 + k[None, :] ** 2)
    f[0, 0] = 1.0 / n
    scr = np.real(np.fft.ifft2(np.fft.fft2(w) * f ** (-11 / 6)))
    scr *= rms_nm * 1e-9 / scr.std()
    return scr


def _median_fps(fn, frames, warmup):
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(frames):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return 1.0 / float(np.median(ts))


def cpu_baseline(validI4Q, opd, frames=5, warmup=2):
    """Reference-faithful numpy restatement (complex128, scipy fft2 / numpy ifft2, thread pool of
    6 or 8 workers over modulation points like the reference's joblib loop) on host cores: the median
    of ``frames`` frames after ``warmup`` warm-up frames."""
    from oracle import pwfs as opw
    ncpu = os.cpu_count() or 1
    n_jobs = 8 if ncpu > 16 else 6
    pup = opw.telescope_pupil(N_RES)
    flux = opw.flux_map(pup, opw.n_photon('I3', 8), TS, DIAM)
    o = opw.PyramidOracle(N_RES, N_SUBAP, MOD, 0.5, pup, flux, 0.8e-6, n_pix_separation=SEP,
                          n_pix_edge=EDGE, nTheta_user_defined=T_POINTS, precision='fp64', n_jobs=n_jobs,
                          validI4Q=validI4Q)
    fps = _median_fps(lambda: o.measure_opd(opd), frames, warmup)
    # the same oracle over every core of the process's affinity set
    n_all = len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else ncpu
    if o._pool is not None:
        o._pool.shutdown()
    o.n_jobs = n_all
    o._pool = None
    fps_all = _median_fps(lambda: o.measure_opd(opd), frames, 1)
    if o._pool is not None:
        o._pool.shutdown()
    # complex64 variant of the same restatement (the reference with precision 32), reference thread count
    o32 = opw.PyramidOracle(N_RES, N_SUBAP, MOD, 0.5, pup, flux, 0.8e-6, n_pix_separation=SEP,
                            n_pix_edge=EDGE, nTheta_user_defined=T_POINTS, precision='fp32', n_jobs=n_jobs,
                            validI4Q=validI4Q)
    fps32 = _median_fps(lambda: o32.measure_opd(opd), frames, 1)
    if o32._pool is not None:
        o32._pool.shutdown()
    model = ''
    try:
        with open('/proc/cpuinfo') as f:
            model = next((ln.split(':', 1)[1].strip() for ln in f if ln.startswith('model name')), '')
    except OSError:
        pass
    env = {k: os.environ.get(k) for k in ('OMP_NUM_THREADS', 'OPENBLAS_NUM_THREADS', 'MKL_NUM_THREADS')}
    return dict(value=fps, unit='frames/s', cores=n_jobs, kind='port',
                sample=f'median of {frames} full config-2 frames ({T_POINTS} points, 2 x 1280^2 complex128 FFTs '
                       f'each) after {warmup} warm-up frames, {n_jobs} threads; host reports {ncpu} CPUs',
                all_cores={'value': fps_all, 'threads': n_all},
                fp32={'value': fps32, 'threads': n_jobs},
                host={'cpu_count': ncpu, 'affinity': n_all, 'model': model, 'thread_env': env})


def closed_loop_leg(tel, ngs, wfs, dev, n_iter=1000, n_modes=1000):
    """Config-2 closed loop as specified: 2-layer atmosphere (r0 0.15 m, L0 25 m, 10 and 20 m/s), 81x81 DM
    (5209 actuators), KL basis (float64 on the GPU), push-pull interaction matrix of the first 1000 KL
    modes at 10 nm through the PWFS, float64 pseudo-inverse, reconstructor M2C @ M, integrator gain 0.4,
    2-frame delay, 1000 iterations with the Strehl ratio every 10 iterations; everything on the device."""
    import torch
    from types import SimpleNamespace
    from oopao_amd import Atmosphere, DeformableMirror, InteractionMatrix
    from oopao_amd.closed_loop import run_closed_loop
    from oopao_amd.calibration.compute_KL_modal_basis import compute_KL_basis
    dm = DeformableMirror(tel, N_SUBAP, 0.35, print_dm_properties=False)
    t0 = time.perf_counter()
    M2C = compute_KL_basis(tel, SimpleNamespace(r0=0.15, L0=25.0), dm)[:, :n_modes]
    torch.cuda.synchronize()
    t_kl = time.perf_counter() - t0
    t0 = time.perf_counter()
    calib = InteractionMatrix(ngs, tel, dm, wfs, M2C, 10e-9, nMeasurements=100, single_pass=False)
    R = (torch.from_numpy(M2C).to(dev) @ calib.M_dev).float().contiguous()
    torch.cuda.synchronize()
    t_im = time.perf_counter() - t0
    atm = Atmosphere(tel, 0.15, 25, windSpeed=[10, 20], fractionalR0=[0.7, 0.3], windDirection=[0, 90],
                     altitude=[0, 10000])
    atm.initializeAtmosphere(tel)
    atm.generateNewPhaseScreen(10)
    run_closed_loop(tel, ngs, atm, dm, wfs, R, gain=0.4, n_iter=5)          # warm-up
    atm.generateNewPhaseScreen(10)           # every timed loop starts from the same screens
    dm.coefs = np.zeros(dm.nValidAct)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = run_closed_loop(tel, ngs, atm, dm, wfs, R, gain=0.4, n_iter=n_iter, strehl_every=10)
    dt = time.perf_counter() - t0
    # the same loop (same screens, same start) with the reconstructor held in BF16 (plain, and split
    # BF16 = hi + lo) on the matrix-core GEMV path, and with the reconstructor in its factored form
    # (M2C, M): M2C @ (M @ s), 61 MB per iteration instead of 210 MB; each residual trace is compared with
    # the float32 loop's
    half = n_iter // 2
    r32 = out['wfe_residual']
    forms = {}
    factors = (M2C, calib.M_dev)
    for key, prec, rec in (('bf16', 'bf16', R), ('bf16x3', 'bf16x3', R), ('factored', 'fp32', factors),
                           ('factored_bf16x3', 'bf16x3', factors)):
        atm.generateNewPhaseScreen(10)
        dm.coefs = np.zeros(dm.nValidAct)
        run_closed_loop(tel, ngs, atm, dm, wfs, rec, gain=0.4, n_iter=5, reconstructor_precision=prec)
        atm.generateNewPhaseScreen(10)
        dm.coefs = np.zeros(dm.nValidAct)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        o = run_closed_loop(tel, ngs, atm, dm, wfs, rec, gain=0.4, n_iter=n_iter, strehl_every=10,
                            reconstructor_precision=prec)
        dtp = time.perf_counter() - t0
        r = o['wfe_residual']
        forms[key] = {'value': round(n_iter / dtp, 2), 'unit': 'iterations/s',
                       'wfe_residual_nm': round(float(np.mean(r[half:])), 1),
                       'strehl_pct_mean_second_half': round(float(np.mean(o['strehl'][half // 10:])), 2),
                       'residual_trace_max_rel_diff_vs_fp32': float(np.max(np.abs(r - r32) / r32)),
                       'residual_trace_max_rel_diff_vs_fp32_first_30': float(np.max(np.abs(r[:30] - r32[:30])
                                                                                 / r32[:30]))}
    bf16 = forms['bf16']
    return {'value': round(n_iter / dt, 2), 'unit': 'iterations/s', 'iterations': n_iter, 'nAct': int(dm.nValidAct),
            'nSignal': int(wfs.nSignal), 'modes': n_modes, 'dm_route': 'dense' if dm.use_dense else 'separable',
            'kl_seconds': round(t_kl, 2), 'im_seconds': round(t_im, 2), 'im_cond': round(float(calib.cond), 1),
            'wfe_atmosphere_nm': round(float(np.mean(out['wfe_atmosphere'][half:])), 1),
            'wfe_residual_nm': round(float(np.mean(out['wfe_residual'][half:])), 1),
            'strehl_pct_mean_second_half': round(float(np.mean(out['strehl'][half // 10:])), 2),
            'note': 'wall clock of the loop incl. atmosphere update (host RNG draws) and Strehl every 10 it.',
            'bf16_reconstructor': bf16, 'split_bf16_reconstructor': forms['bf16x3'],
            'factored_reconstructor': dict(forms['factored'], form='M2C @ (M @ s), float32'),
            'factored_split_bf16_reconstructor': dict(forms['factored_bf16x3'], form='M2C @ (M @ s), split BF16')}


def atmosphere_leg(tel, n_steps=200):
    """Config-2 atmosphere updates (2 layers, 10 and 20 m/s at 1 kHz: 0.4 / 0.8 px per step, screens
    with 326^2 supports): frozen flow and boiling layers (t_boiling 5 and 10 ms), with the reference's
    host RandomState draws and with the device counter-based draws; wall clock per update incl. the
    summed OPD (steps/s)."""
    import torch
    from oopao_amd import Atmosphere
    out = {}
    for name, t_b, rng in (('frozen_host_rng', None, 'host'), ('frozen_device_rng', None, 'device'),
                           ('boiling_host_rng', [0.005, 0.01], 'host'), ('boiling_device_rng', [0.005, 0.01], 'device')):
        atm = Atmosphere(tel, 0.15, 25, windSpeed=[10, 20], fractionalR0=[0.7, 0.3], windDirection=[0, 90],
                         altitude=[0, 10000], t_boiling=t_b, rng=rng)
        atm.initializeAtmosphere(tel)
        atm.generateNewPhaseScreen(10)
        for _ in range(5):
            atm.update()
        torch.cuda.synchronize()
        steps = n_steps if rng == 'device' or t_b is None else n_steps // 4
        t0 = time.perf_counter()
        for _ in range(steps):
            atm.update()
        torch.cuda.synchronize()
        out[name] = {'steps_per_s': round(steps / (time.perf_counter() - t0), 1), 'steps': steps}
    return out


def _wall_ms(fn, reps=10):
    import torch
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps * 1e3


def extended_leg(dev):
    """Wall-clock timings of the sensors and propagation modes beyond the Pyramid frame, at the config-1
    pupil (80 px, 8 m): the Bi-O Edge WFS (OOPAO's tutorial sensor: 20 x 20 subapertures, unmodulated,
    grey width 3, and a modulated one: 3 lambda/D, 48 points), the gain-sensing camera (calibration and one
    optical-gain estimate on 100 DM modes through the config-1 PWFS's full focal camera), an asterism of 5
    sources (two off-axis NGS, two sodium LGS, one on axis) through a 2-layer atmosphere with a 60" field,
    and the same atmosphere with angular-spectrum propagation (scintillation) for 3 sources; and an ELT-scale
    PWFS frame."""
    import torch
    from oopao_amd import Telescope, Source, Atmosphere, Asterism, DeformableMirror, Pyramid
    from oopao_amd.BioEdge import BioEdge
    from oopao_amd.GainSensingCamera import GainSensingCamera
    out = {}
    tel = Telescope(80, 8.0, 1e-3)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    opd = np.random.RandomState(3).randn(80, 80) * 30e-9
    tel.OPD_no_pupil = opd
    for name, kw in (('bioedge_tutorial', dict(modulation=0, grey_width=3, lightRatio=0.01, n_pix_separation=0,
                                               postProcessing='fullFrame_incidence_flux')),
                     ('bioedge_modulated', dict(modulation=3, grey_width=1, lightRatio=0.1, n_pix_separation=20,
                                                nTheta_user_defined=48, postProcessing='slopesMaps'))):
        wfs = BioEdge(nSubap=20, telescope=tel, **kw)

        def relay(wfs=wfs):
            tel.OPD_no_pupil = opd
            ngs * tel * wfs
        out[name] = {'ms_per_frame': round(_wall_ms(relay), 3), 'N': int(wfs.resolution),
                     'nTheta': int(wfs.nTheta), 'nSignal': int(wfs.nSignal)}
    dm = DeformableMirror(tel, 20, 0.35, print_dm_properties=False)
    pwfs = Pyramid(20, tel, 3, 0.5, n_pix_separation=20, n_pix_edge=10, nTheta_user_defined=48)
    N = pwfs.nRes
    pwfs.focal_plane_camera.resolution = N
    basis = dm.modes.reshape(80, 80, dm.nValidAct)[:, :, :100] * tel.pupil[:, :, None]
    t0 = time.perf_counter()
    gsc = GainSensingCamera(pwfs.mask, basis)
    tel.resetOPD()
    ngs * tel * pwfs
    pwfs * pwfs.focal_plane_camera
    pwfs.focal_plane_camera * gsc
    torch.cuda.synchronize()
    t_cal = time.perf_counter() - t0

    def og():
        tel.OPD_no_pupil = opd
        ngs * tel * pwfs
        pwfs * pwfs.focal_plane_camera
        pwfs.focal_plane_camera * gsc
    out['gain_sensing_camera'] = {'calibration_ms': round(t_cal * 1e3, 2),
                                  'ms_per_estimate_incl_frame_and_focal_camera': round(_wall_ms(og), 3),
                                  'modes': 100, 'N': int(N)}
    telf = Telescope(80, 8.0, 1e-3, fov=60)
    ngs2 = Source('I3', 8, display_properties=False)
    ngs2 * telf
    for name, kw, srcs in (
            ('asterism_offaxis_lgs', {}, [Source('I3', 8, coordinates=[0, 0], display_properties=False),
                                          Source('I3', 8, coordinates=[20, 35], display_properties=False),
                                          Source('I3', 8, coordinates=[25, 215], display_properties=False),
                                          Source('Na', 8, coordinates=[0, 0], altitude=90e3, display_properties=False),
                                          Source('Na', 8, coordinates=[15, 200], altitude=90e3,
                                                 display_properties=False)]),
            ('scintillation_asm', {'angular_spectrum_propagation': True},
             [Source('I3', 8, coordinates=[0, 0], display_properties=False),
              Source('I3', 8, coordinates=[8, 60], display_properties=False),
              Source('Na', 8, coordinates=[0, 0], altitude=90e3, display_properties=False)])):
        atm = Atmosphere(telf, 0.15, 25, windSpeed=[10, 20], fractionalR0=[0.7, 0.3], windDirection=[0, 90],
                         altitude=[0, 10000], **kw)
        atm.initializeAtmosphere(telf)
        atm.generateNewPhaseScreen(12)
        ast = Asterism(srcs)

        def step(atm=atm, ast=ast):
            atm.update()
            ast ** atm * telf
        out[name] = {'ms_per_update_and_relay': round(_wall_ms(step), 3), 'sources': len(srcs), 'n': 80}
    # ELT-scale PWFS: 39 m pupil, 30 % obstruction, 480 px, 120 x 120 subapertures at 4 px (the 100 x 100 at
    # 4.8 px variant is not constructible, Pyramid.py:216-217), N = 1920, 96 points (32-lane transforms)
    elt = Telescope(480, 39.0, 1e-3, centralObstruction=0.3)
    ngs3 = Source('I3', 8, display_properties=False)
    ngs3 * elt
    wfs_elt = Pyramid(120, elt, 3, 0.5, n_pix_separation=120, n_pix_edge=60, nTheta_user_defined=96)
    opd_elt = np.random.RandomState(5).randn(480, 480) * 50e-9

    def elt_frame():
        elt.OPD_no_pupil = opd_elt
        ngs3 * elt * wfs_elt
    out['elt_pwfs_120x120'] = {'ms_per_frame': round(_wall_ms(elt_frame, reps=5), 2), 'N': int(wfs_elt.nRes),
                               'nTheta': int(wfs_elt.nTheta), 'nSignal': int(wfs_elt.nSignal),
                               'path': wfs_elt.engine.mode}
    # the same telescope sampled at 512 px with 128 x 128 subapertures (N = 2048, register-resident path)
    elt2 = Telescope(512, 39.0, 1e-3, centralObstruction=0.3)
    ngs4 = Source('I3', 8, display_properties=False)
    ngs4 * elt2
    wfs_elt2 = Pyramid(128, elt2, 3, 0.5, n_pix_separation=128, n_pix_edge=64, nTheta_user_defined=96)
    opd_elt2 = np.random.RandomState(6).randn(512, 512) * 50e-9

    def elt2_frame():
        elt2.OPD_no_pupil = opd_elt2
        ngs4 * elt2 * wfs_elt2
    out['elt_pwfs_128x128'] = {'ms_per_frame': round(_wall_ms(elt2_frame, reps=10), 3), 'N': int(wfs_elt2.nRes),
                               'nTheta': int(wfs_elt2.nTheta), 'nSignal': int(wfs_elt2.nSignal),
                               'path': wfs_elt2.engine.mode}
    # the config-5 sensor as specified: 100 x 100 subapertures at 4.8 px on the 480 px pupil (unbinned
    # frame on the separable path, detector integration by area)
    wfs_elt3 = Pyramid(100, elt, 3, 0.5, n_pix_separation=100, n_pix_edge=50, nTheta_user_defined=96,
                       fractional_sampling=True)

    def elt3_frame():
        elt.OPD_no_pupil = opd_elt
        ngs3 * elt * wfs_elt3
    out['elt_pwfs_100x100_4p8px'] = {'ms_per_frame': round(_wall_ms(elt3_frame, reps=5), 3),
                                     'N': int(wfs_elt3.nRes), 'cam': int(wfs_elt3.cam.resolution),
                                     'nTheta': int(wfs_elt3.nTheta), 'nSignal': int(wfs_elt3.nSignal),
                                     'path': 'separable, area-integrated detector'}
    return out


def im_leg(dev, n_modes=1000):
    """Interaction-matrix modes/s: push-pull of n_modes seeded random DM modes (+-10 nm) through the
    config-1 PWFS (80 px pupil, 20x20 subapertures, 48 points), batched, FP64 SVD included."""
    import torch
    from oopao_amd import Telescope, Source, DeformableMirror, Pyramid, InteractionMatrix
    tel = Telescope(80, 8.0, 1e-3)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    dm = DeformableMirror(tel, 20, 0.35, print_dm_properties=False)
    wfs = Pyramid(20, tel, 3, 0.5, n_pix_separation=20, n_pix_edge=10, nTheta_user_defined=48)
    M2C = np.random.RandomState(1).randn(dm.nValidAct, n_modes)
    M2C /= np.std(M2C, axis=0)
    # warm-up on the timed call's routes: the same batch (workspace size) and >= 256 modes, so that the
    # float64 pseudo-inverse takes the Gram-eigensolver route of the timed 632 x 1000 matrix
    InteractionMatrix(ngs, tel, dm, wfs, M2C[:, :300], 10e-9, nMeasurements=250, single_pass=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    calib = InteractionMatrix(ngs, tel, dm, wfs, M2C, 10e-9, nMeasurements=250, single_pass=False)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {'value': round(n_modes / dt, 1), 'unit': 'modes/s (push-pull, incl. FP64 SVD)', 'modes': n_modes,
            'nSignal': int(wfs.nSignal), 'cond': float(calib.cond)}


def im_batched_leg(n_modes=1000):
    """Push-pull calibration of four PWFS systems of uneven sampling in one HIP graph: the config-1
    sensor (80 px, 20x20 subapertures, 48 points) and 40 / 64 / 80 subapertures at 4 px each with
    48 / 72 / 96 points, each with its own Cartesian DM (nSubap + 1 actuators across) and KL basis
    (n_modes modes, or nAct - 1 when fewer), stroke 10 nm; tile and batch sizes from the autotuner;
    float64 SVD pseudo-inverses on the GPU timed separately."""
    import torch
    from types import SimpleNamespace
    from oopao_amd import Telescope, Source, DeformableMirror, Pyramid
    from oopao_amd.calibration.compute_KL_modal_basis import compute_KL_basis
    from oopao_amd.calibration.batched import BatchedCalibration
    specs = [(80, 20, 48), (160, 40, 48), (256, 64, 72), (320, 80, 96)]
    systems = []
    for n, ns, T in specs:
        tel = Telescope(n, 8.0, 1e-3)
        ngs = Source('I3', 8, display_properties=False)
        ngs * tel
        dm = DeformableMirror(tel, ns, 0.35, print_dm_properties=False)
        wfs = Pyramid(ns, tel, 3, 0.5, n_pix_separation=ns, n_pix_edge=ns // 2, nTheta_user_defined=T)
        kl = compute_KL_basis(tel, SimpleNamespace(r0=0.15, L0=25.0), dm)
        M2C = kl[:, :min(n_modes, kl.shape[1])]
        systems.append((ngs, tel, dm, wfs, M2C))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bc = BatchedCalibration(systems, 10e-9, tune=True)
    torch.cuda.synchronize()
    t_tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    bc.run_eager()
    torch.cuda.synchronize()
    t_eager = time.perf_counter() - t0
    D_eager = [D.clone() for D in bc.run_eager()]

    def graph_time(concurrent):
        bc.capture(concurrent=concurrent)
        bc.run()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        Ds = bc.run()
        e1.record()
        torch.cuda.synchronize()
        same = all(torch.equal(D, E) for D, E in zip(Ds, D_eager))
        return e0.elapsed_time(e1) / 1e3, same
    t_serial, same_serial = graph_time(False)
    t_graph, same_graph = graph_time(True)
    t0 = time.perf_counter()
    vaults = bc.vaults()
    torch.cuda.synchronize()
    t_svd = time.perf_counter() - t0
    modes = sum(cs.nModes for cs in bc.systems)
    per = [{'n': cs.tel.resolution, 'nSubap': cs.wfs.nSubap, 'nTheta': cs.wfs.nTheta, 'nAct': cs.dm.nValidAct,
            'modes': cs.nModes, 'nSignal': int(cs.wfs.nSignal), 'batch_modes': cs.batch,
            'autotune': {k: v for k, v in t.items() if k != 'trials'}, 'cond': float(cv.cond)}
           for cs, t, cv in zip(bc.systems, bc.tuning, vaults)]
    return {'value': round(modes / t_graph, 1), 'unit': 'modes/s (push-pull, 4 sensors, one graph launch)',
            'modes': modes, 'graph_s': round(t_graph, 4), 'eager_s': round(t_eager, 4),
            'graph_serial_s': round(t_serial, 4), 'graph_speedup_vs_eager': round(t_eager / t_graph, 3),
            'graph_equals_eager': bool(same_graph and same_serial),
            'graph_note': 'one branch per system in the graph (forked streams); graph_serial_s records the '
                          'systems one after another',
            'autotune_s': round(t_tune, 2), 'svd_pinv_f64_s': round(t_svd, 3), 'systems': per}


def adjoint_leg(n_screens=64, steps=10):
    """Differentiable PWFS (config 1 sensor: 80 px pupil, 20x20 subapertures, 48 points): gradients of
    ||s(phi) - s_target||^2 w.r.t. 64 seeded phase screens per step through the hand-written adjoint
    (forward + backward, modulation recomputed in backward)."""
    import torch
    from oopao_amd import Telescope, Source, Pyramid
    from oopao_amd.differentiable import pwfs_signal
    tel = Telescope(80, 8.0, 1e-3)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    wfs = Pyramid(20, tel, 3, 0.5, n_pix_separation=20, n_pix_edge=10, nTheta_user_defined=48)
    rs = np.random.RandomState(100)
    ph = torch.from_numpy((rs.randn(n_screens, 80, 80) * 0.3 * tel.pupil).astype(np.float32)).cuda()
    ph.requires_grad_(True)
    tgt = torch.zeros((1, wfs.nSignal), device='cuda')
    for _ in range(3):
        ph.grad = None
        ((pwfs_signal(wfs, ph).double() - tgt) ** 2).sum().backward()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ph.grad = None
        ((pwfs_signal(wfs, ph).double() - tgt) ** 2).sum().backward()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    # Jacobians through the adjoint at an aberrated working point: the full slopes Jacobian for the
    # 357 DM influence functions (one VJP per slope), and optical gains of those modes (one VJP per mode)
    from oopao_amd import DeformableMirror
    from oopao_amd.differentiable import slopes_jacobian, optical_gains_adjoint
    dm = DeformableMirror(tel, 20, 0.35, print_dm_properties=False)
    pupil = torch.from_numpy(tel.pupil.astype(np.float32)).cuda()
    modes = dm._build_modes().t().reshape(-1, 80, 80) * pupil
    zero = torch.zeros((80, 80), device='cuda')
    slopes_jacobian(wfs, zero, modes[:8], batch=64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    D0 = slopes_jacobian(wfs, zero, modes, batch=64)
    torch.cuda.synchronize()
    t_jac = time.perf_counter() - t0
    t0 = time.perf_counter()
    og = optical_gains_adjoint(wfs, ph[0].detach(), modes, D0, batch=64)
    torch.cuda.synchronize()
    t_og = time.perf_counter() - t0
    return {'value': round(n_screens / dt, 1), 'unit': 'gradients/s', 'screens_per_step': n_screens,
            'ms_per_step': round(dt * 1e3, 3), 'note': 'forward + adjoint, float32, wall clock',
            'jacobian': {'nSignal': int(wfs.nSignal), 'modes': int(modes.shape[0]), 'seconds': round(t_jac, 4)},
            'optical_gains': {'modes': int(modes.shape[0]), 'seconds': round(t_og, 4),
                              'mean_og': round(float(og.mean()), 4)}}


def kl_leg(tel):
    """KL modal basis of the config-2 DM (81 x 81 grid, 5209 actuators) in float64 on the GPU:
    IF spectra, HHt, seed basis and KL diagonalisation (compute_KL_basis defaults)."""
    import torch
    from types import SimpleNamespace
    from oopao_amd import DeformableMirror
    from oopao_amd.calibration.compute_KL_modal_basis import compute_KL_basis
    dm = DeformableMirror(tel, N_SUBAP, 0.35, print_dm_properties=False)
    atm = SimpleNamespace(r0=0.15, L0=25.0)
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats()
    t0 = time.perf_counter()
    B = compute_KL_basis(tel, atm, dm)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {'seconds': round(dt, 3), 'nAct': int(dm.nValidAct), 'modes': int(B.shape[1]), 'dtype': 'f64',
            'peak_mem_GB': round(torch.cuda.max_memory_allocated() / 1e9, 2)}


def _event_time(f, n):
    import torch
    f()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(n):
        f()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / n * 1e-3


def _rel(got, exact):
    d = (got.double() - exact)
    return {'rel_l2': float(d.norm() / exact.norm()),
            'max_rel': float(d.abs().max() / exact.abs().max())}


def contraction_leg(tel, dev, n_modes=1000):
    """Dense influence-function and reconstructor contractions in BF16 on the matrix cores
    (FP32 accumulation), each with its error against the float64 product.

    * config-2 DM (81 x 81 grid, 5209 actuators, 320 x 320 pupil): OPD of n_modes unit-RMS seeded
      commands (the batched IM / KL-mode OPDs) as one GEMM [102400 x 5209] x [5209 x n_modes], BF16 and
      split BF16, beside the float32 GEMV route (one matrix read per command);
    * ELT scale: 39 m pupil, 480 x 480, 30 % obstruction, 80 x 80 DM (BF16 IF [230400 x nAct]) applied to
      one command (GEMV), and a BF16 reconstructor MVM [nAct x 15700] against the float32 GEMV."""
    import ctypes
    import torch
    from oopao_amd import Telescope, DeformableMirror, _lib
    from oopao_amd._device import ptr, stream_ptr
    from oopao_amd.contract import Contraction, PackedBF16
    out = {}
    dm = DeformableMirror(tel, N_SUBAP, 0.35, print_dm_properties=False)
    modes = dm._build_modes()
    R, K = modes.shape
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    X = torch.randn(K, n_modes, device=dev, generator=g) * 20e-9
    C = Contraction(PackedBF16(modes, split=True))
    opd = torch.empty((n_modes, R), dtype=torch.float32, device=dev)
    cols = torch.arange(0, n_modes, n_modes // 8, device=dev)
    exact = (modes.double() @ X[:, cols].double()).t()
    res = {'R': R, 'K': K, 'B': n_modes}
    for prec, terms in (('bf16', 1), ('bf16x3', 3)):
        t = _event_time(lambda: C.gemm(X, precision=prec, out=opd), 10)
        tf = 2.0 * R * K * n_modes / t / 1e12
        res[prec] = {'ms': round(t * 1e3, 4), 'TFLOP_s': round(tf, 1),
                     'mfma_TFLOP_s_issued': round(tf * terms, 1),
                     'frac_bf16_peak': round(tf * terms / BF16_MFMA_PEAK_TFLOPS, 4),
                     'err_vs_f64': _rel(opd[cols], exact)}
    y = torch.empty(R, dtype=torch.float32, device=dev)
    x0 = X[:, 0].contiguous()
    t = _event_time(lambda: _lib.call('oopao_gemv', ptr(modes), ptr(x0), ptr(y), None, R, K, 1.0, 0.0,
                                      stream_ptr(dev)), 20)
    res['fp32_gemv_per_command_ms'] = round(t * 1e3, 4)
    res['fp32_gemv_GBs'] = round(R * K * 4 / t / 1e9, 1)
    out['config2_dm_gemm'] = res
    del C, opd, exact, modes
    torch.cuda.empty_cache()

    elt = Telescope(480, 39.0, 1e-3, centralObstruction=0.3)
    dm5 = DeformableMirror(elt, N_SUBAP, 0.35, print_dm_properties=False)
    modes = dm5._build_modes()
    R, K = modes.shape
    C = Contraction(PackedBF16(modes, split=False))
    x = (torch.randn(K, device=dev, generator=g) * 20e-9).contiguous()
    y = torch.empty(R, dtype=torch.float32, device=dev)
    t16 = _event_time(lambda: C.gemv(x, out=y), 20)
    err = _rel(y, modes.double() @ x.double())
    y32 = torch.empty(R, dtype=torch.float32, device=dev)
    t32 = _event_time(lambda: _lib.call('oopao_gemv', ptr(modes), ptr(x), ptr(y32), None, R, K, 1.0, 0.0,
                                        stream_ptr(dev)), 20)
    out['elt_dm_gemv'] = {'R': R, 'K': K, 'bf16_ms': round(t16 * 1e3, 4),
                          'bf16_GBs': round(C.A.hi.numel() * 2 / t16 / 1e9, 1), 'bf16_err_vs_f64': err,
                          'fp32_ms': round(t32 * 1e3, 4), 'fp32_GBs': round(R * K * 4 / t32 / 1e9, 1),
                          'fp32_err_vs_f64': _rel(y32, modes.double() @ x.double())}
    del C, modes
    torch.cuda.empty_cache()
    nsig = 15700
    Rm = torch.randn(K, nsig, device=dev, generator=g) * 1e-3
    s = torch.randn(nsig, device=dev, generator=g)
    C = Contraction(PackedBF16(Rm, split=False))
    c = torch.empty(K, dtype=torch.float32, device=dev)
    t16 = _event_time(lambda: C.gemv(s, out=c), 50)
    ex = Rm.double() @ s.double()
    c32 = torch.empty(K, dtype=torch.float32, device=dev)
    t32 = _event_time(lambda: _lib.call('oopao_gemv', ptr(Rm), ptr(s), ptr(c32), None, K, nsig, 1.0, 0.0,
                                        stream_ptr(dev)), 50)
    out['elt_reconstructor_mvm'] = {'R': K, 'K': nsig, 'bf16_us': round(t16 * 1e6, 2),
                                    'bf16_GBs': round(C.A.hi.numel() * 2 / t16 / 1e9, 1),
                                    'bf16_err_vs_f64': _rel(c, ex), 'fp32_us': round(t32 * 1e6, 2),
                                    'fp32_GBs': round(K * nsig * 4 / t32 / 1e9, 1),
                                    'fp32_err_vs_f64': _rel(c32, ex)}
    del C, Rm
    torch.cuda.empty_cache()
    return out


def measured_traffic(kernels, workload):
    """HBM bytes per launch of the given kernels of one workload ('slopes_only' or 'full_frame') from
    the committed rocprofv3 --pmc summary of this bench (FETCH_SIZE and WRITE_SIZE in KiB per dispatch,
    scaled by the copy-kernel calibration)."""
    path = os.path.join(ROOT, 'profiles', 'pwfs_config2_traffic.json')
    try:
        with open(path) as f:
            d = json.load(f)
        tot = 0.0
        for k in kernels:
            e = d['workloads'][workload][k]
            tot += (e['fetch_kib'] * d['fetch_scale'] + e['write_kib'] * d['write_scale']) * 1024
        return int(tot)
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--settle-ms', type=float, default=100.0,
                    help='before the warm-up steps of each timed loop, repeat its step for this long (untimed) so '
                         'that the GPU clock has left its idle state: after 10 steps (~2 ms) alone the config-2 '
                         'frame measures ~3 %% slower than after 200')
    ap.add_argument('--cpu-frames', type=int, default=5)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--batch', type=int, default=8, help='screens per call for the batched form')
    ap.add_argument('--no-extra', action='store_true',
                    help='skip the closed-loop, interaction-matrix, KL-basis, adjoint and contraction legs')
    ap.add_argument('--legs', default='closed_loop,interaction_matrix,im_batched,kl_basis,adjoint,contractions,extended,'
                                      'atmosphere',
                    help='comma-separated extra legs to run')
    ap.add_argument('--calib-copy', action='store_true',
                    help='also run one elementwise copy of 256 MiB (counter calibration for rocprof --pmc)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py: this build measures one GPU only (no multi-GPU path); run with --gpus 1',
              file=sys.stderr)
        sys.exit(2)

    import torch
    from oopao_amd import Telescope, Source, Pyramid

    dev = torch.device('cuda', 0)
    torch.cuda.set_device(dev)
    if args.calib_copy:
        x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)   # 256 MiB
        y = x.mul(1.0)
        torch.cuda.synchronize()
        del x, y
    tel = Telescope(N_RES, DIAM, TS, centralObstruction=0.0)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    wfs = Pyramid(N_SUBAP, tel, MOD, 0.5, n_pix_separation=SEP, n_pix_edge=EDGE, nTheta_user_defined=T_POINTS)
    eng, se = wfs.engine, wfs.slopes_engine
    opd_host = synthetic_opd(N_RES) * tel.pupil
    opd = torch.from_numpy(opd_host.astype(np.float32)).to(dev)[None].contiguous()
    ngs ** tel
    amp = wfs._amp()
    scale = 2 * np.pi / ngs.wavelength
    buf = eng.buffers(1)
    frame = torch.empty((1, eng.cam, eng.cam), dtype=torch.float32, device=dev)
    sig2d = torch.empty((1, 2 * se.w, se.w), dtype=torch.float32, device=dev)
    sig = torch.empty((1, se.nsig), dtype=torch.float32, device=dev)
    # the slopes read only the quadrant columns of the detector (Pyramid.py:838-853, 946-961)
    bands = wfs.slope_bands()
    n_cols = sum(b - a for a, b in bands)

    def make_step(bd):
        def step(ev=None):
            if ev is not None:
                ev[0].record()
            eng.rows(opd, scale, amp, buf, bd)
            if ev is not None:
                ev[1].record()
            eng.cols(buf, frame, bd)
            if ev is not None:
                ev[2].record()
            se.slopes(frame, sig2d, sig)
        return step

    def settle(step):
        """Repeat ``step`` (untimed) for args.settle_ms of wall clock."""
        t_end = time.perf_counter() + args.settle_ms * 1e-3
        while time.perf_counter() < t_end:
            for _ in range(10):
                step()
            torch.cuda.synchronize()

    def timed(step, K):
        """ms per step over K steps (bracketed by synchronisations) and the mean rows / columns
        pass durations from HIP events on the kernels' stream."""
        settle(step)
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(K)]
        t0 = time.perf_counter()
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record()
        for k in range(K):
            step(evs[k])
        end.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return (start.elapsed_time(end) / K, float(np.mean([e[0].elapsed_time(e[1]) for e in evs])),
                float(np.mean([e[1].elapsed_time(e[2]) for e in evs])), wall)

    K = args.steps
    step = make_step(bands)
    ms_step, rows_ms, cols_ms, wall = timed(step, K)
    fps = 1000.0 / ms_step
    sig_slopes_only = sig.clone()
    # the whole frame: every detector column (same slopes bit for bit)
    step_full = make_step(None)
    ms_full, rows_full, cols_full, _ = timed(step_full, K)
    full_equal = bool(torch.equal(sig, sig_slopes_only))

    # algorithmic work of the dominant kernel (columns pass + chunk reduction), per launch
    n, N, T, P = N_RES, eng.N, eng.T, eng.P
    log2N = np.log2(N)

    def cols_work(ncol):
        # one pruned forward + one inverse N-point FFT per pupil-plane column (4 per detector column) and
        # point; bytes: the intermediate's columns read once, the chunk partials written and read by the
        # reduction, the mask vector and the frame written
        flops = 4 * ncol * T * 2 * 5 * N * log2N
        nbytes = T * n * 4 * ncol * 8 + buf['nchunk'] * eng.cam * ncol * 4 * 2 + N * 8 + eng.cam ** 2 * 4
        return flops, nbytes
    cols_flops, cols_bytes = cols_work(n_cols)
    achieved_tf = cols_flops / (cols_ms * 1e-3) / 1e12
    achieved_gbs = cols_bytes / (cols_ms * 1e-3) / 1e9
    full_flops, full_bytes = cols_work(eng.cam)
    full_tf = full_flops / (cols_full * 1e-3) / 1e12

    # batched form: B screens per call (slopes only)
    B = args.batch
    opdB = opd.expand(B, N_RES, N_RES).contiguous()
    bufB = eng.buffers(B)
    frameB = torch.empty((B, eng.cam, eng.cam), dtype=torch.float32, device=dev)
    sig2dB = torch.empty((B, 2 * se.w, se.w), dtype=torch.float32, device=dev)
    sigB = torch.empty((B, se.nsig), dtype=torch.float32, device=dev)

    def stepB():
        eng.rows(opdB, scale, amp, bufB, bands)
        eng.cols(bufB, frameB, bands)
        se.slopes(frameB, sig2dB, sigB)
    settle(stepB)
    for _ in range(3):
        stepB()
    torch.cuda.synchronize()
    s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    KB = max(3, K // B)
    s0.record()
    for _ in range(KB):
        stepB()
    s1.record()
    torch.cuda.synchronize()
    batched_fps = B * KB * 1000.0 / s0.elapsed_time(s1)

    # the single-screen slopes-only frame captured once in a HIP graph (the library's calls are
    # stream-ordered and allocate nothing) and replayed K times; its slopes must equal the eager ones
    # bit for bit
    step()
    torch.cuda.synchronize()
    sig_eager = sig.clone()
    gs = torch.cuda.Stream(dev)
    gs.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(gs):
        for _ in range(3):
            step()
    torch.cuda.current_stream(dev).wait_stream(gs)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step()
    for _ in range(5):
        graph.replay()
    torch.cuda.synchronize()
    graph_match = bool(torch.equal(sig, sig_eager))
    g0 = torch.cuda.Event(enable_timing=True)
    g1 = torch.cuda.Event(enable_timing=True)
    g0.record()
    for _ in range(K):
        graph.replay()
    g1.record()
    torch.cuda.synchronize()
    graph_ms = g0.elapsed_time(g1) / K
    graph_match = graph_match and bool(torch.equal(sig, sig_eager))
    del graph

    extra = {}
    if not args.no_extra:
        legs = set(args.legs.split(','))
        if 'closed_loop' in legs:
            extra['closed_loop'] = closed_loop_leg(tel, ngs, wfs, dev)
        if 'interaction_matrix' in legs:
            extra['interaction_matrix'] = im_leg(dev)
        if 'im_batched' in legs:
            extra['im_batched'] = im_batched_leg()
        if 'kl_basis' in legs:
            extra['kl_basis'] = kl_leg(tel)
        if 'adjoint' in legs:
            extra['adjoint'] = adjoint_leg()
        if 'contractions' in legs:
            extra['contractions'] = contraction_leg(tel, dev)
        if 'extended' in legs:
            extra['extended'] = extended_leg(dev)
        if 'atmosphere' in legs:
            extra['atmosphere'] = atmosphere_leg(tel)

    out = {
        'metric': 'PWFS frames/s at 80x80 subaps, 72 mod pts',
        'value': round(fps, 2),
        'unit': 'frames/s',
        'n_gpus': 1,
        'steps': K,
        'warmup': args.warmup,
        'settle_ms': args.settle_ms,
        'ms_per_step': round(ms_step, 5),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'fp32/complex64',
        'data': 'synthetic seeded OPD screen (150 nm RMS, seed 10), resident in HBM',
        'config': {'workload': 'config-2 modulated PWFS, OPD -> slopesMaps signal: n=320, N=1280, T=72, the %d '
                               'detector columns of the quadrants computed (slopes-only frame; the whole 320x320 '
                               'frame is full_frame)' % n_cols,
                   'n': N_RES, 'N': int(N), 'nTheta': int(T), 'nSubap': N_SUBAP, 'batch': 1,
                   'nchunk': buf['nchunk'], 'pts_per_task': buf['ppt'], 'tip_tables': int(P),
                   'column_bands': [list(b) for b in bands]},
        'kernels_ms': {'rows_pass': round(rows_ms, 5), 'cols_pass_plus_reduce': round(cols_ms, 5)},
        'roofline': {'bound': 'valu', 'kernel': 'k_cols_sep<1280,64> + k_reduce_chunks (slopes-only columns)',
                     'achieved': round(achieved_tf, 3), 'peak': FP32_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': round(achieved_tf / FP32_PEAK_TFLOPS, 4),
                     'frac_of_nonpacked_peak': round(achieved_tf / VALU_PEAK_TFLOPS, 4),
                     'peak_note': 'peak = published MI355X FP32 157.3 TFLOP/s; the build disables packed FP32, '
                                  'whose non-packed FMA ceiling is 78.6 TFLOP/s',
                     'traffic': measured_traffic(['k_cols_sep<1280,64>', 'k_reduce_chunks'], 'slopes_only'),
                     'traffic_source': 'profiles/pwfs_config2_traffic.json (rocprofv3 --pmc FETCH_SIZE / '
                                       'WRITE_SIZE of this bench, copy-calibrated)',
                     'flops_per_launch': cols_flops, 'bytes_per_launch': cols_bytes,
                     'hbm_achieved_GBs': round(achieved_gbs, 1),
                     'hbm_frac': round(achieved_gbs / HBM_PEAK_GBS, 4)},
        'full_frame': {'value': round(1000.0 / ms_full, 2), 'unit': 'frames/s', 'ms_per_step': round(ms_full, 5),
                       'kernels_ms': {'rows_pass': round(rows_full, 5), 'cols_pass_plus_reduce': round(cols_full, 5)},
                       'cols_achieved_TFLOPs': round(full_tf, 3),
                       'cols_frac': round(full_tf / FP32_PEAK_TFLOPS, 4),
                       'traffic': measured_traffic(['k_cols_sep<1280,64>', 'k_reduce_chunks'], 'full_frame'),
                       'slopes_equal_slopes_only': full_equal},
        'batched': {'batch': B, 'value': round(batched_fps, 2), 'unit': 'frames/s'},
        'graph': {'value': round(1000.0 / graph_ms, 2), 'unit': 'frames/s', 'ms_per_step': round(graph_ms, 5),
                  'replays': K, 'slopes_equal_eager': graph_match},
        'wall_s_timed': round(wall, 4),
    }
    out.update(extra)
    if not args.no_cpu:
        valid = wfs.validI4Q.copy()
        out['cpu_baseline'] = cpu_baseline(valid, opd_host, frames=args.cpu_frames)
        out['speedup_vs_cpu'] = round(fps / out['cpu_baseline']['value'], 1)
    print(json.dumps(out))


if __name__ == '__main__':
    main()
