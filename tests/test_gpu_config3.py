"""GPU: config 3 at its workload -- push-pull calibration of 1000 KL modes (10 nm) through the config-1
PWFS and three more sensors of uneven sampling (40, 64 and 80 subapertures at 4 px with 48, 72 and 96
modulation points), each with its own 8 m Cartesian DM and KL basis, all recorded into one HIP graph
(oopao_amd/calibration/batched.py; calibration/InteractionMatrix.py:85-158 semantics).

Checks: the graph replay equals the eager run bit for bit; six sampled columns per system (first, second,
eighth, 101st, middle and last KL mode) equal the float64 oracle's push-pull columns
0.5 (s(+a) - s(-a)) / a within 1e-4 rel-L2 (oracle/pwfs.PyramidOracle on 16 threads, DM OPD through
oracle/dm.DMOracle's pinned separable identity).
"""
import numpy as np
import pytest
import torch

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

SPECS = [(80, 20, 48), (160, 40, 48), (256, 64, 72), (320, 80, 96)]


def test_config3_batched_calibration_vs_oracle():
    from types import SimpleNamespace
    from oopao_amd import Telescope, Source, DeformableMirror, Pyramid
    from oopao_amd.calibration.compute_KL_modal_basis import compute_KL_basis
    from oopao_amd.calibration.batched import BatchedCalibration
    from oracle import pwfs as opw
    from oracle.dm import DMOracle
    stroke = 10e-9
    systems = []
    for n, ns, T in SPECS:
        print('config 3: building n=%d nSubap=%d T=%d' % (n, ns, T), flush=True)
        tel = Telescope(n, 8.0, 1e-3)
        ngs = Source('I3', 8, display_properties=False)
        ngs * tel
        dm = DeformableMirror(tel, ns, 0.35, print_dm_properties=False)
        wfs = Pyramid(ns, tel, 3, 0.5, n_pix_separation=ns, n_pix_edge=ns // 2, nTheta_user_defined=T)
        kl = compute_KL_basis(tel, SimpleNamespace(r0=0.15, L0=25.0), dm)
        M2C = np.asarray(kl[:, :min(1000, kl.shape[1])])
        systems.append((ngs, tel, dm, wfs, M2C))
    bc = BatchedCalibration(systems, stroke, tune=True)
    eager = [D.clone() for D in bc.run_eager()]
    Ds = bc.run()
    torch.cuda.synchronize()
    for D, E in zip(Ds, eager):
        assert torch.equal(D, E)
    for (ngs, tel, dm, wfs, M2C), (n, ns, T), D in zip(systems, SPECS, Ds):
        nm = M2C.shape[1]
        assert D.shape == (wfs.nSignal, nm)
        flux = opw.flux_map(tel.pupil, ngs.nPhoton, tel.samplingTime, tel.D)
        o = opw.PyramidOracle(n, ns, 3, 0.5, tel.pupil, flux, ngs.wavelength, n_pix_separation=ns,
                              n_pix_edge=ns // 2, nTheta_user_defined=T, n_jobs=16)
        od = DMOracle(n, 8.0, ns, 0.35)
        try:
            assert np.array_equal(wfs.validI4Q, o.validI4Q)
            worst = 0.0
            for m in (0, 1, 7, 100, nm // 2, nm - 1):
                opd = od.opd_separable(M2C[:, m] * stroke)
                col = 0.5 * (o.measure_opd(opd)[1] - o.measure_opd(-opd)[1]) / stroke
                worst = max(worst, rel_l2(D[:, m].double().cpu().numpy(), col))
            print('n=%d nSubap=%d T=%d: %d modes, nSignal %d, worst sampled column rel-L2 vs float64 oracle %.2e'
                  % (n, ns, T, nm, wfs.nSignal, worst), flush=True)
            assert worst < 1e-4
        finally:
            if o._pool is not None:
                o._pool.shutdown()
