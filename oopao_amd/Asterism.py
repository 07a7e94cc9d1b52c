"""Asterism (Asterism.py:12-263): a list of sources propagated together.

``ast ** atm * tel * wfs`` propagates every source: the atmosphere computes each source's
footprint OPD (off-axis / laser guide stars, Atmosphere.fill_OPD_support), the telescope and
the wavefront sensor relay each source in turn (their relay methods loop over ``ast.src``).
Per-source quantities are exposed as lists, as in the reference.
"""
import numpy as np

from .tools import OopaoError


class Asterism:
    def __init__(self, list_src):
        if not list_src:
            raise OopaoError('an Asterism needs at least one source')
        self.n_source = len(list_src)
        self.src = list_src
        for i, src in enumerate(self.src):
            src.inAsterism = True
            src.ast_idx = i
        self.chromatic_shift = None
        self.tag = 'asterism'
        self.type = 'asterism'
        self.wavelength = self.src[0].wavelength
        print(self)

    def _each(self, name):
        return [getattr(s, name) for s in self.src]

    @property
    def laser_coordinates(self):
        return self._each('laser_coordinates')

    @property
    def Na_profile(self):
        return self._each('Na_profile')

    @property
    def fwhm_spot_up(self):
        return self._each('fwhm_spot_up')

    @property
    def fluxMap(self):
        return self._each('fluxMap')

    @property
    def phase(self):
        return self._each('phase')

    @phase.setter
    def phase(self, val):
        for s in self.src:
            s.phase = val[s.ast_idx]

    @property
    def phase_no_pupil(self):
        return self._each('phase_no_pupil')

    @property
    def coordinates(self):
        return self._each('coordinates')

    @property
    def altitude(self):
        return self._each('altitude')

    @property
    def nPhoton(self):
        return self._each('nPhoton')

    @property
    def OPD(self):
        return np.array(self._each('OPD'))

    @property
    def OPD_no_pupil(self):
        return np.array(self._each('OPD_no_pupil'))

    @property
    def scintillation(self):
        return np.array(self._each('scintillation'), dtype=object)

    @scintillation.setter
    def scintillation(self, val):
        for s in self.src:
            s.scintillation = val[s.ast_idx]

    @property
    def scintillation_no_pupil(self):
        return np.array(self._each('scintillation_no_pupil'), dtype=object)

    @property
    def intensity(self):
        return self._each('intensity')

    @intensity.setter
    def intensity(self, val):
        for s in self.src:
            s.intensity = val[s.ast_idx]

    def __pow__(self, obj):
        """Re-propagation (Asterism.py:200-207)."""
        obj.src = self
        for s in self.src:
            s.optical_path = [[s.type + '(' + s.optBand + ')', s]]
        self.reset()
        self * obj
        return self

    def __mul__(self, obj):
        """Propagation (Asterism.py:209-217)."""
        if obj is not None:
            if isinstance(obj, list):
                for o in obj:
                    self * o
            else:
                obj.relay(self)
        return self

    def reset(self):
        for s in self.src:
            s.reset()

    def print_optical_path(self):
        for s in self.src:
            s.print_optical_path()

    def print_properties(self):
        print(self)

    def __repr__(self):
        rows = ['%-6s|%-10s|%-10s|%-12s' % ('Source', 'Zenith["]', 'Azimuth[°]', 'Altitude[m]')]
        for i, s in enumerate(self.src):
            rows.append('%-6d|%-10.2f|%-10.2f|%-12.1f' % (i, s.coordinates[0], s.coordinates[1], s.altitude))
        w = max(len(r) for r in rows)
        return '\n'.join([f'{" Asterism ":-^{w}}'] + rows + ['-' * w]) + '\n'
