"""GPU: off-axis and laser-guide-star sources through the atmosphere (oopao_amd.Atmosphere with a
field of view, oopao_amd.Asterism) against the float64 oracle (oracle/atmosphere.source_layer).

Reference: set_pupil_footprint / relay / fill_OPD_support (Atmosphere.py:423-445, 533-622), the
bilinear interpolations of tools/interpolateGeometricalTransformation.py:152-265, Asterism.py.
The oracle takes each layer's shifted screen (layer.OPD) from the device (its bicubic shift and clip
are pinned by tests/test_gpu_atmosphere.py) and applies the footprint, the sub-pixel translation
and the cone magnification in float64. Tolerance: per-source OPD rel-L2 <= 1e-6.
"""
import numpy as np
import pytest

from oracle.atmosphere import source_layer
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def _system():
    from oopao_amd import Telescope, Source, Atmosphere
    tel = Telescope(32, 8, 1e-3, fov=60)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    atm = Atmosphere(tel, 0.15, 25, windSpeed=[10, 20], fractionalR0=[0.7, 0.3], windDirection=[0, 90],
                     altitude=[0, 10000])
    atm.initializeAtmosphere(tel)
    atm.generateNewPhaseScreen(11)
    return tel, ngs, atm


def _expected(atm, tel, s):
    out = np.zeros((tel.resolution, tel.resolution))
    for i in range(atm.nLayer):
        lay = getattr(atm, 'layer_' + str(i + 1))
        ch = 0.0 if s.chromatic_shift is None else s.chromatic_shift[i]
        sub = source_layer(atm.layer_opd(i), tel.resolution, lay.D, lay.altitude, s.coordinates[0], s.coordinates[1],
                           s.altitude, ch)
        out += sub * atm.wavelength / 2 / np.pi * np.sqrt(atm.fractionalR0[i])
    return out


def test_asterism_footprints_and_cone_effect():
    from oopao_amd import Source, Asterism
    from oopao_amd.tools import OopaoError
    print('off-axis / LGS asterism', flush=True)
    tel, ngs, atm = _system()
    srcs = [Source('I3', 8, coordinates=[0, 0], display_properties=False),
            Source('I3', 8, coordinates=[20, 35], display_properties=False),
            Source('Na', 8, coordinates=[0, 0], altitude=90e3, display_properties=False),
            Source('Na', 8, coordinates=[15, 200], altitude=90e3, display_properties=False),
            Source('I3', 8, coordinates=[10, 120], chromatic_shift=[0.0, 1.7], display_properties=False)]
    ast = Asterism(srcs)
    for step in range(3):
        if step:
            atm.update()
        ast ** atm
        for k, s in enumerate(srcs):
            ref = _expected(atm, tel, s)
            err = rel_l2(s.OPD_no_pupil, ref)
            print('step %d source %d (zenith %.0f", az %.0f, alt %g) rel-L2 %.2e' % (step, k, s.coordinates[0],
                                                                                s.coordinates[1], s.altitude, err))
            assert err < 1e-6
        assert atm.OPD.shape == (5, 32, 32)
    # the on-axis member equals the single on-axis propagation
    ngs ** atm
    assert rel_l2(srcs[0].OPD_no_pupil, ngs.OPD_no_pupil) < 1e-7
    # a source outside the field of view is refused
    with pytest.raises(OopaoError):
        Source('I3', 8, coordinates=[40, 0], display_properties=False) ** atm


def test_asterism_through_pyramid():
    from oopao_amd import Source, Asterism, Pyramid
    print('asterism through the PWFS', flush=True)
    tel, ngs, atm = _system()
    wfs = Pyramid(8, tel, 2, 0.1, n_pix_separation=8, n_pix_edge=4, nTheta_user_defined=16)
    srcs = [Source('I3', 8, coordinates=[0, 0], display_properties=False),
            Source('I3', 8, coordinates=[20, 35], display_properties=False)]
    ast = Asterism(srcs)
    ast ** atm * tel * wfs
    sig = np.asarray(wfs.signal)
    assert sig.shape == (2, wfs.nSignal)
    for k, s in enumerate(srcs):
        s ** atm * tel * wfs
        assert rel_l2(sig[k], np.asarray(wfs.signal)) < 1e-6


@pytest.mark.parametrize('mode', ['wrapped', 'geometric', 'unwrap'])
def test_scintillation_vs_oracle(mode):
    """Angular-spectrum propagation (Atmosphere.py:644-706, 938-972) against oracle.atmosphere.scintillation_field:
    intensity and phasor rel-L2 <= 1e-4 (complex64 FFTs), geometric backup OPD <= 1e-6."""
    from oopao_amd import Telescope, Source, Atmosphere, Asterism
    from oracle.atmosphere import scintillation_field
    print('scintillation', mode, flush=True)
    tel = Telescope(64, 8, 1e-3, fov=20)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    atm = Atmosphere(tel, 0.15, 25, windSpeed=[10, 20], fractionalR0=[0.7, 0.3], windDirection=[0, 90],
                     altitude=[0, 10000], angular_spectrum_propagation=True,
                     geometric_phase_backup=(mode == 'geometric'), unwrap_diffractive_phase=(mode == 'unwrap'))
    atm.initializeAtmosphere(tel)
    atm.generateNewPhaseScreen(12)
    srcs = [Source('I3', 8, coordinates=[0, 0], display_properties=False),
            Source('I3', 8, coordinates=[8, 60], display_properties=False),
            Source('Na', 8, coordinates=[0, 0], altitude=90e3, display_properties=False)]
    ast = Asterism(srcs)
    ast ** atm
    order = sorted(range(atm.nLayer), key=lambda i: getattr(atm, 'layer_' + str(i + 1)).altitude, reverse=True)
    for k, s in enumerate(srcs):
        lops, alts = [], []
        for i in order:
            lay = getattr(atm, 'layer_' + str(i + 1))
            w = atm.wavelength / 2 / np.pi * np.sqrt(atm.fractionalR0[i])
            lops.append(source_layer(atm.layer_opd(i), 64, lay.D, lay.altitude, s.coordinates[0], s.coordinates[1],
                                     s.altitude) * w)
            alts.append(lay.altitude)
        E = scintillation_field(lops, alts, s.wavelength, tel.pixelSize)
        kk = 2 * np.pi / s.wavelength
        ei = rel_l2(s.scintillation_no_pupil, np.abs(E) ** 2)
        opd = s.OPD_no_pupil
        geo = sum(lops)
        if mode == 'geometric':
            ep = rel_l2(opd, geo)
            assert ep < 1e-6
        elif mode == 'wrapped':
            ep = rel_l2(np.exp(1j * kk * opd), E / np.abs(E))
            assert ep < 1e-4
        else:
            ref = geo + np.angle(E * np.exp(-1j * kk * geo)) / kk
            ep = rel_l2(opd, ref)
            assert ep < 1e-4
        print('source %d intensity %.2e phase %.2e' % (k, ei, ep))
        assert ei < 1e-4
    assert atm.scintillation_map.shape == (3, 64, 64)
