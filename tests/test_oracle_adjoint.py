"""CPU: pins of the float64 PWFS adjoint oracle (oracle/adjoint.py) by central finite differences of
the oracle forward (Pyramid.py:612-646, 712-757, 835-854) along random directions."""
import numpy as np

from oracle import adjoint as oadj
from tests.helpers import small_system, random_phase


def test_frame_vjp_matches_finite_differences():
    o = small_system()
    rs = np.random.RandomState(3)
    ph = random_phase(o.geom.n, 4, 1.0, o.pupil)
    G = rs.randn(o.geom.cam, o.geom.cam)
    grad = oadj.frame_vjp(o, ph, G)
    eps = 1e-5
    for k in range(3):
        d = rs.randn(*ph.shape) * o.pupil
        Lp = np.sum(G * o.frame(ph + eps * d))
        Lm = np.sum(G * o.frame(ph - eps * d))
        fd = (Lp - Lm) / (2 * eps)
        an = np.sum(grad * d)
        assert abs(fd - an) <= 1e-7 * abs(fd), (fd, an)
    # outside the pupil the field is zero, so is the gradient
    assert np.max(np.abs(grad[o.pupil == 0])) == 0


def test_slopes_vjp_matches_finite_differences():
    o = small_system()
    rs = np.random.RandomState(5)
    ph = random_phase(o.geom.n, 6, 1.0, o.pupil)
    frame = o.frame(ph)
    dsig = rs.randn(o.nSignal)
    gF = oadj.slopes_vjp(o, frame, dsig)
    from oracle.pwfs import slopes_maps
    eps = 1e-6 * frame.max()
    for k in range(3):
        d = rs.randn(*frame.shape)
        sp = slopes_maps(frame + eps * d, o.validI4Q, o.ref2d, o.geom)[1]
        sm = slopes_maps(frame - eps * d, o.validI4Q, o.ref2d, o.geom)[1]
        fd = np.sum(dsig * (sp - sm)) / (2 * eps)
        an = np.sum(gF * d)
        assert abs(fd - an) <= 1e-6 * max(abs(fd), 1e-12), (fd, an)
