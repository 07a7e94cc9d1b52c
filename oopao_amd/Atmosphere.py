"""Atmosphere: multi-layer frozen-flow von Karman turbulence (Atmosphere.py:26-1246).

Initialisation follows the reference on the host in float64 (phase screens
from the same RandomState seeds, Assemat extrusion matrices A, B from the von
Karman covariances; Atmosphere.py:225-374, 707-810, 1017-1036). The per-step
update runs on the GPU (oopao_amd/csrc/atm.hip): whole-pixel shifts and ring
extrusion (add_row), the clipped cubic sub-pixel shift and the pupil-footprint
sum of the layers. The extrusion noise xi is drawn from each layer's host
RandomState in the reference's order, so a seeded run reproduces the
reference's screens.

Sources: on-axis and off-axis natural guide stars and laser guide stars at a
finite altitude, single or in an Asterism, with a field of view (layers sized
for it, Atmosphere.py:297-300). Each source's OPD follows fill_OPD_support
(Atmosphere.py:589-622): the footprint of its line of sight on each layer
(set_pupil_footprint, 423-445), the sub-pixel remainder of the footprint
position applied as a bilinear translation of the layer, and for a laser guide
star the cone-effect magnification (bilinear, about the footprint centre),
on the GPU (oopao_bilinear_affine). With angular_spectrum_propagation the source
field is propagated through the layers by the angular spectrum (oopao_asm_propagate,
Atmosphere.py:644-706, 938-972), giving the scintillation (intensity) and the phase
(wrapped, geometric, or geometric guide + wrapped residual). Boiling layers
(t_boiling, Atmosphere.py:376-408) mix their extended screen with a fresh
phase screen each step on the GPU (oopao_atm_boil), from the layer's host
RandomState(boiling_seed) draws in the reference's order.
"""
import ctypes

import numpy as np
import torch
from numpy.random import RandomState

from . import _lib
from ._device import ptr, stream_ptr, to_host
from .phaseStats import (ft_sh_phase_screen, ft_phase_screen, makeCovarianceMatrix, extrusion_rings,
                         extrusion_operators, centred_dft_matrix, ft_amplitude, subharmonic_terms,
                         subharmonic_table)
from .tools import OopaoError, warning


class _Layer:
    pass


class Atmosphere:
    def __init__(self, telescope, r0: float, L0: float, windSpeed: list, fractionalR0: list, windDirection: list,
                 altitude: list, src=None, param=None, elevation: float = 90.0, mode: float = 2,
                 angular_spectrum_propagation: bool = False, geometric_phase_backup: bool = False,
                 unwrap_diffractive_phase=False, rytov_var: float = None, rytov_wvl: float = 500e-9,
                 t_boiling=None, rng: str = 'host'):
        self.tag = 'atmosphere'
        self.device = telescope.device
        # where the per-step draws come from: 'host' = each layer's numpy RandomState in the reference's
        # order (a seeded run reproduces the reference's screens); 'device' = counter-based Philox4x32-10
        # draws on the GPU keyed by the same seeds (same statistics, other numbers, no host work per step)
        if rng not in ('host', 'device'):
            raise OopaoError("rng must be 'host' or 'device', got %r" % (rng,))
        self.rng = rng
        self.hasNotBeenInitialized = True
        self.angular_spectrum_propagation = bool(angular_spectrum_propagation)
        self.geometric_phase_backup = geometric_phase_backup
        self.unwrap_diffractive_phase = unwrap_diffractive_phase
        self.rytov_wvl = rytov_wvl
        self.sampling_checked = False
        self._saturation_warned = False
        self._transfer = {}
        self.altitude = altitude
        self.wavelength = 500 * 1e-9
        self._elevation = max(5.0, float(elevation))
        el = np.radians(self._elevation)
        self.altitude_zenith = [h * np.sin(el) for h in self.altitude]
        self.r0_def = 0.15
        self.rad2arcsec = (180. / np.pi) * 3600
        self.telescope = telescope
        self.r0 = r0
        self.fractionalR0 = fractionalR0
        self.cn2 = (self.r0 ** (-5. / 3) / (0.423 * (2 * np.pi / self.wavelength) ** 2)) / np.max([1, np.max(self.altitude)])
        self.L0 = L0
        self.nLayer = len(fractionalR0)
        self.windSpeed = windSpeed
        self.windDirection = windDirection
        self.t_boiling = t_boiling
        self.n_extra_pixel = 2
        self.V0 = (np.sum(np.asarray(self.fractionalR0) * np.asarray(self.windSpeed) ** (5 / 3))) ** (3 / 5)
        self.tau0 = 0.31 * self.r0 / self.V0
        self.is_user_defined_opd = False
        self.mode = mode
        self.seeingArcsec = self.rad2arcsec * (self.wavelength / self.r0)
        if src is None and self.telescope.src is None:
            raise OopaoError('The Atmosphere object requires a Source. Either provide a Source directly as an '
                             'attribute, or propagate the Source through the Telescope before creating the Atmosphere.')
        self.src = src if src else self.telescope.src
        self.src_list = [self.src] if self.src.tag == 'source' else self.src.src
        self.asterism = None if self.src.tag == 'source' else self.src
        self.param = param
        self._OPD_dev = None
        self.update_rytov_variance()
        if rytov_var is not None:
            self.rytov_var = float(rytov_var)

    # ------------------------------------------------------------ init (host)
    def initializeAtmosphere(self, telescope=None, compute_covariance=True):
        """Atmosphere.initializeAtmosphere (Atmosphere.py:225-279)."""
        if telescope is not None:
            self.telescope = telescope
        self.compute_covariance = compute_covariance
        self.fov = self.telescope.fov
        self.fov_rad = self.telescope.fov_rad
        if self.hasNotBeenInitialized:
            self.initial_r0 = self.r0
            for i in range(self.nLayer):
                print('Creation of layer' + str(i + 1) + '/' + str(self.nLayer) + ' ...')
                lay = self.buildLayer(self.telescope, self.r0_def, self.L0, i_layer=i)
                setattr(self, 'layer_' + str(i + 1), lay)
        else:
            print('Re-setting the atmosphere to its initial state...')
            self.r0 = self.initial_r0
        self.hasNotBeenInitialized = False
        self.r0 = self.r0                       # recomputes B (Atmosphere.py:275, 1027-1035)
        self.generateNewPhaseScreen(0)
        self.relay(self.src)
        print(self)

    def buildLayer(self, telescope, r0, L0, i_layer):
        """Atmosphere.buildLayer (Atmosphere.py:281-374)."""
        tel = self.telescope
        lay = _Layer()
        lay.randomState = RandomState(42 + i_layer * 1000)
        lay.altitude = self.altitude[i_layer]
        lay.windSpeed = self.windSpeed[i_layer]
        lay.direction = self.windDirection[i_layer]
        lay.vY = lay.windSpeed * np.cos(np.deg2rad(lay.direction))
        lay.vX = lay.windSpeed * np.sin(np.deg2rad(lay.direction))
        lay.D_fov = tel.D + 2 * np.tan(self.fov_rad / 2) * lay.altitude
        lay.resolution_fov = int(np.ceil((tel.resolution / tel.D) * lay.D_fov))
        lay.resolution = lay.resolution_fov + 4
        lay.center = lay.resolution // 2
        lay.D = lay.resolution * tel.D / tel.resolution
        lay.pixel_size = lay.D / lay.resolution
        lay.n_extra_pixel = self.n_extra_pixel
        lay.t_boiling = self.t_boiling[i_layer]
        lay.boiling_seed = 100000 + i_layer * 100000      # Atmosphere.py:329-333
        lay.rng_key, lay.rng_draw = (42 + i_layer * 1000) & 0xFFFFFFFF, 0
        res, M = lay.resolution, lay.resolution + self.n_extra_pixel
        lay.OPD_host = ft_sh_phase_screen(self.r0, self.L0, res, lay.D / res, seed=i_layer)
        lay.seed = i_layer
        outer, inner, lay.innerZ, lay.outerZ = extrusion_rings(res, self.n_extra_pixel, lay.D)
        lay.outerMask, lay.innerMask = outer, inner
        # covariance matrices (Atmosphere.py:707-777); the ring geometry depends only on the layer
        # resolution, so layers of equal size share them
        cache = self.__dict__.setdefault('_cov_cache', {})
        key = (res, self.L0)
        if key not in cache:
            print('Computing the covariance matrices ...')
            ZZt = makeCovarianceMatrix(lay.innerZ, lay.innerZ, self.L0, self.r0_def)
            cache[key] = (ZZt, makeCovarianceMatrix(lay.innerZ, lay.outerZ, self.L0, self.r0_def),
                          makeCovarianceMatrix(lay.outerZ, lay.outerZ, self.L0, self.r0_def), np.linalg.pinv(ZZt))
        self.ZZt, self.ZXt, self.XXt, self.ZZt_inv = cache[key]
        lay.ZZt, lay.ZXt, lay.XXt, lay.ZZt_inv = self.ZZt, self.ZXt, self.XXt, self.ZZt_inv
        lay.A, lay.B = extrusion_operators(lay.ZZt, lay.ZXt, lay.XXt, lay.ZZt_inv, self.r0_def, self.r0)
        lay.notDoneOnce = True
        lay.buff = np.zeros(2)
        lay.ratio = np.zeros(2)
        # device state
        dev = self.device
        lay.M = M
        lay.inner_idx = torch.from_numpy(np.flatnonzero(inner.reshape(-1) != 0).astype(np.int32)).to(dev)
        lay.outer_idx = torch.from_numpy(np.flatnonzero(outer.reshape(-1) != 0).astype(np.int32)).to(dev)
        lay.A_dev = torch.from_numpy(np.ascontiguousarray(lay.A)).to(dev)
        lay.map = torch.zeros(M * M, dtype=torch.float64, device=dev)
        lay.tmp = torch.zeros(M * M, dtype=torch.float64, device=dev)
        lay.z = torch.zeros(lay.inner_idx.numel(), dtype=torch.float64, device=dev)
        lay.X = torch.zeros(lay.outer_idx.numel(), dtype=torch.float64, device=dev)
        lay.mm = torch.zeros(4, dtype=torch.float64, device=dev)   # (min, max) + working space
        lay.xi = torch.zeros(lay.outer_idx.numel(), dtype=torch.float64, device=dev)
        return lay

    def _set_B(self, lay):
        lay.B_dev = torch.from_numpy(np.ascontiguousarray(lay.B)).to(self.device)

    def save_state(self, path):
        """Write the evolving state of every layer to an .npz file: the extruded screen, its (min, max),
        the sub-pixel shift buffer, the per-step ratio, the layer's RandomState and boiling seed, and the
        parameters load_state checks. The reference has no
        simulation checkpoint; a second initializeAtmosphere resets the layers (Atmosphere.py:246-271).
        load_state on an atmosphere built with the same parameters resumes the same sequence of screens."""
        if self.hasNotBeenInitialized:
            raise OopaoError('The Atmosphere object needs to be initialised using the initialiseAtmosphere()')
        out = {'nLayer': np.int64(self.nLayer), 'r0': np.float64(self.r0), 'L0': np.float64(self.L0),
               'resolution': np.int64(self.telescope.resolution),
               'samplingTime': np.float64(self.telescope.samplingTime), 'rng': np.str_(self.rng)}
        for i in range(self.nLayer):
            lay = getattr(self, 'layer_' + str(i + 1))
            key, pos, has_gauss, cached = lay.randomState.get_state()[1:]
            p = 'layer_%d_' % (i + 1)
            out[p + 'M'] = np.int64(lay.M)
            out[p + 'map'] = lay.map.cpu().numpy()
            out[p + 'mm'] = lay.mm[:2].cpu().numpy()
            out[p + 'buff'] = np.asarray(lay.buff, np.float64)
            out[p + 'ratio'] = np.asarray(lay.ratio, np.float64)
            out[p + 'notDoneOnce'] = np.bool_(lay.notDoneOnce)
            out[p + 'shift'] = np.asarray(getattr(lay, 'shift', (0.0, 0.0)), np.float64)
            out[p + 'rs_key'] = np.asarray(key, np.uint32)
            out[p + 'rs_pos'] = np.int64(pos)
            out[p + 'rs_gauss'] = np.asarray([has_gauss, cached], np.float64)
            for name, val in self._layer_params(i).items():
                out[p + name] = np.float64(val)
            out[p + 'boiling_seed'] = np.int64(lay.boiling_seed)
            out[p + 'rng_key'] = np.int64(lay.rng_key)
            out[p + 'rng_draw'] = np.int64(lay.rng_draw)
        np.savez(path, **out)

    def _layer_params(self, i):
        """Parameters that fix layer i's per-step evolution (wind shift per step, statistics, boiling)."""
        lay = getattr(self, 'layer_' + str(i + 1))
        t_b = getattr(lay, 't_boiling', None)
        return {'vX': lay.vX, 'vY': lay.vY, 'pixel_size': lay.pixel_size, 'altitude': lay.altitude,
                'fractionalR0': self.fractionalR0[i], 't_boiling': np.nan if t_b is None else t_b}

    def load_state(self, path):
        """Restore what save_state wrote. The atmosphere must have the saved one's parameters: number of
        layers, resolution, r0, L0, sampling time, and per layer the wind (speed and direction), pixel
        size, altitude, fractional r0 and boiling time; otherwise the resumed screens would follow other
        statistics, and OopaoError names what differs."""
        if self.hasNotBeenInitialized:
            raise OopaoError('The Atmosphere object needs to be initialised using the initialiseAtmosphere()')
        with np.load(path) as f:
            if int(f['nLayer']) != self.nLayer or int(f['resolution']) != self.telescope.resolution \
                    or float(f['r0']) != float(self.r0) or float(f['L0']) != float(self.L0):
                raise OopaoError('the saved atmosphere state does not match this atmosphere (nLayer, resolution, '
                                 'r0, L0)')
            if 'rng' in f and str(f['rng']) != self.rng:
                raise OopaoError("the saved atmosphere state used rng=%r, this atmosphere rng=%r"
                                 % (str(f['rng']), self.rng))
            if 'samplingTime' in f and float(f['samplingTime']) != float(self.telescope.samplingTime):
                raise OopaoError('the saved atmosphere state was taken at another sampling time (%g s, here %g s)'
                                 % (float(f['samplingTime']), self.telescope.samplingTime))
            for i in range(self.nLayer):
                p = 'layer_%d_' % (i + 1)
                for name, val in self._layer_params(i).items():
                    if p + name not in f:
                        continue
                    saved = float(f[p + name])
                    same = (np.isnan(saved) and np.isnan(val)) or (not np.isnan(saved) and not np.isnan(val)
                                                                  and np.isclose(saved, val, rtol=1e-12, atol=0))
                    if not same:
                        raise OopaoError('the saved atmosphere state does not match layer %d: %s saved %r, here %r'
                                         % (i + 1, name, saved, float(val)))
            for i in range(self.nLayer):
                lay = getattr(self, 'layer_' + str(i + 1))
                p = 'layer_%d_' % (i + 1)
                if int(f[p + 'M']) != lay.M:
                    raise OopaoError('the saved atmosphere state does not match layer %d (M)' % (i + 1))
                if p + 'boiling_seed' in f:
                    lay.boiling_seed = int(f[p + 'boiling_seed'])
                if p + 'rng_key' in f:
                    lay.rng_key, lay.rng_draw = int(f[p + 'rng_key']), int(f[p + 'rng_draw'])
                lay.map.copy_(torch.from_numpy(f[p + 'map']))
                lay.mm[:2].copy_(torch.from_numpy(f[p + 'mm']))
                lay.buff = f[p + 'buff'].copy()
                lay.ratio = f[p + 'ratio'].copy()
                lay.notDoneOnce = bool(f[p + 'notDoneOnce'])
                lay.shift = tuple(float(v) for v in f[p + 'shift'])
                g = f[p + 'rs_gauss']
                lay.randomState = RandomState()
                lay.randomState.set_state(('MT19937', f[p + 'rs_key'], int(f[p + 'rs_pos']), int(g[0]), float(g[1])))
        self._compute_opd()
        if self.telescope.isPaired:
            self.telescope.src ** self * self.telescope

    def generateNewPhaseScreen(self, seed=None):
        """Atmosphere.generateNewPhaseScreen (Atmosphere.py:779-810)."""
        if seed is None:
            import time
            t = time.localtime()
            seed = t.tm_hour * 3600 + t.tm_min * 60 + t.tm_sec
        for i in range(self.nLayer):
            lay = getattr(self, 'layer_' + str(i + 1))
            if self.mode == 2:
                phase = ft_sh_phase_screen(self.r0, self.L0, lay.resolution, lay.D / lay.resolution, seed=seed + i)
            else:
                phase = ft_phase_screen(self.r0, self.L0, lay.resolution, lay.D / lay.resolution, seed=seed + i)
            lay.OPD_host = phase
            lay.randomState = RandomState(seed + i * 1000)
            Z = phase[lay.innerMask[1:-1, 1:-1] != 0]
            X = np.matmul(lay.A, Z) + np.matmul(lay.B, lay.randomState.normal(size=lay.B.shape[1]))
            ms = np.zeros((lay.M, lay.M))
            ms[lay.outerMask != 0] = X
            ms[lay.outerMask == 0] = phase.reshape(lay.resolution * lay.resolution)
            lay.map.copy_(torch.from_numpy(ms.reshape(-1)))
            _lib.call('oopao_atm_minmax', ptr(lay.map), lay.M * lay.M, ptr(lay.mm), stream_ptr(self.device))
            lay.notDoneOnce = True
            lay.buff = np.zeros(2)
            lay.shift = (0.0, 0.0)
            lay.boiling_seed = 100000 + seed + i * 100000      # Atmosphere.py:797-799
            lay.rng_key, lay.rng_draw = (seed + i * 1000) & 0xFFFFFFFF, 0
        self._compute_opd()
        if self.telescope.isPaired:
            self.telescope.src ** self * self.telescope

    # ----------------------------------------------------------- device update
    def _add_row(self, lay, step):
        """Atmosphere.add_row (Atmosphere.py:410-421) on the device."""
        sx, sy = int(step[0]), int(step[1])
        if self.rng == 'device':
            # draw lay.rng_draw of the layer's extrusion stream (key (seed + 1000 i, 0))
            _lib.call('oopao_normal_philox', ptr(lay.xi), lay.outer_idx.numel(), lay.rng_key, 0, 0, lay.rng_draw,
                      stream_ptr(self.device))
            lay.rng_draw += 1
        else:
            self._stage_noise(lay, lay.randomState.normal(size=lay.outer_idx.numel()))
        _lib.call('oopao_atm_add_row', ptr(lay.map), ptr(lay.tmp), lay.M, sx, sy, ptr(lay.inner_idx),
                  lay.inner_idx.numel(), ptr(lay.outer_idx), lay.outer_idx.numel(), ptr(lay.A_dev), ptr(lay.B_dev),
                  ptr(lay.xi), ptr(lay.z), ptr(lay.X), ptr(lay.mm), stream_ptr(self.device))
        lay.map, lay.tmp = lay.tmp, lay.map

    _NOISE_RING = 8

    def _stage_noise(self, lay, draw):
        self._stage(lay, 'xi', draw, lay.xi)

    def _stage(self, lay, name, draw, dst):
        """Host draw -> device vector ``dst`` without stalling the host: the draw is written into one of a
        ring of pinned buffers (one ring per layer and name) and copied asynchronously on the stream; a
        buffer is rewritten only after the event of its previous copy has completed."""
        rings = lay.__dict__.setdefault('_rings', {})
        r = rings.get(name)
        if r is None or r['buf'][0].numel() != draw.size:
            r = dict(buf=[torch.empty(draw.size, dtype=torch.float64, pin_memory=True)
                          for _ in range(self._NOISE_RING)], ev=[None] * self._NOISE_RING, k=0)
            rings[name] = r
        k = r['k']
        r['k'] = (k + 1) % self._NOISE_RING
        if r['ev'][k] is not None:
            r['ev'][k].synchronize()
        r['buf'][k].numpy()[:] = draw
        dst.copy_(r['buf'][k], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self.device))
        r['ev'][k] = ev

    # ------------------------------------------------------------ boiling
    def get_t_boiling(self, val):
        """Atmosphere.get_t_boiling (Atmosphere.py:376-385): one boiling time (or None) per layer."""
        if val is None:
            return [None] * self.nLayer
        if np.isscalar(val):
            return [val] * self.nLayer
        val = list(val)
        if len(val) != self.nLayer:
            raise OopaoError('Wrong value for t_boiling! Provide either a scalar '
                             '(applied to each layer) or a list with a value per layer.')
        return val

    @property
    def t_boiling(self):
        return self._t_boiling

    @t_boiling.setter
    def t_boiling(self, val):
        self._t_boiling = self.get_t_boiling(val)
        if self.hasNotBeenInitialized is False:
            print('Updating the boiling time-constant(s)...')
            for i in range(self.nLayer):
                getattr(self, 'layer_' + str(i + 1)).t_boiling = self._t_boiling[i]

    def _boil_tables(self, lay):
        """Device tables of the boiling screen on the layer's (res + 2)^2 support: the FFT-grid amplitudes
        sqrt(PSD) df, the centred DFT matrix, the subharmonic grid coordinates and the workspace."""
        M, pixel = lay.M, lay.D / lay.resolution
        key = (M, float(pixel), float(self.r0), float(self.L0))
        cache = self.__dict__.setdefault('_boil_cache', {})
        t = cache.get(key)
        if t is None:
            dev = self.device
            nb = ctypes.c_size_t(0)
            _lib.call('oopao_atm_boil_bytes', M, ctypes.byref(nb))
            t = dict(amp=torch.from_numpy(ft_amplitude(self.r0, self.L0, M, pixel).reshape(-1)).to(dev),
                     W=torch.from_numpy(centred_dft_matrix(M)).to(dev),
                     xc=torch.from_numpy(np.arange(-M / 2, M / 2) * pixel).to(dev),
                     ws=torch.empty(nb.value, dtype=torch.uint8, device=dev),
                     g=torch.empty(2 * M * M + 48, dtype=torch.float64, device=dev),
                     table=torch.from_numpy(subharmonic_table(self.r0, self.L0, M, pixel).reshape(-1)).to(dev))
            cache[key] = t
        return t

    def _apply_boiling(self, lay):
        """Atmosphere.apply_boiling (Atmosphere.py:387-408) on the device: mapShift = alpha mapShift +
        beta phi with alpha = exp(-dt / t_boiling), beta = sqrt(1 - alpha^2) and phi the screen of seed
        boiling_seed + 1 (ft_sh_phase_screen in mode 2, else ft_phase_screen) on the (res + 2)^2 support."""
        t_b = getattr(lay, 't_boiling', None)
        if t_b is None or not np.isfinite(t_b) or t_b <= 0:
            return False
        alpha = float(np.exp(-self.telescope.samplingTime / t_b))
        if alpha >= 1.0 - 1e-6:
            return False
        beta = float(np.sqrt(max(0.0, 1.0 - alpha ** 2)))
        lay.boiling_seed += 1
        M = lay.M
        tb = self._boil_tables(lay)
        # one draw of the FFT-grid coefficients: real parts then imaginary parts (phaseStats.py:192); the
        # subharmonics restart the same seed (a second RandomState, phaseStats.py:225), i.e. reuse the
        # first 54 of these numbers
        nsub = 12 if self.mode == 2 else 0
        dst = tb['g']
        if self.rng == 'device':
            # the screen's draws from the Philox key (boiling_seed, 1); the subharmonics from its first 54
            _lib.call('oopao_normal_philox', ptr(dst), 2 * M * M, lay.boiling_seed & 0xFFFFFFFF, 1, 0, 0,
                      stream_ptr(self.device))
            if nsub:
                _lib.call('oopao_atm_boil_subharmonics', ptr(dst), 2 * M * M, ptr(tb['table']), nsub,
                          ctypes.c_void_p(dst.data_ptr() + 8 * 2 * M * M), stream_ptr(self.device))
        else:
            g = RandomState(lay.boiling_seed).normal(size=2 * M * M)
            staged = g
            if nsub:
                sub, _ = subharmonic_terms(self.r0, self.L0, M, lay.D / lay.resolution, g[:54])
                staged = np.concatenate((g, sub.reshape(-1)))
            self._stage(lay, 'boil', staged, dst[:staged.size])
        _lib.call('oopao_atm_boil', ptr(lay.map), M, ptr(dst), ptr(tb['amp']), ptr(tb['W']),
                  ctypes.c_void_p(dst.data_ptr() + 8 * 2 * M * M) if nsub else None, nsub, ptr(tb['xc']),
                  ctypes.c_double(alpha), ctypes.c_double(beta), ptr(tb['ws']), ctypes.c_size_t(tb['ws'].numel()),
                  stream_ptr(self.device))
        _lib.call('oopao_atm_minmax', ptr(lay.map), M * M, ptr(lay.mm), stream_ptr(self.device))
        return True

    def updateLayer(self, lay, shift=None):
        """Atmosphere.updateLayer (Atmosphere.py:447-516); boiling first (Atmosphere.py:454-457), so the
        decorrelation accumulates in the support and composes with the wind shift."""
        self._apply_boiling(lay)
        if lay.vX == 0 and lay.vY == 0 and shift is None:
            return      # no wind: the (boiled) support's interior, read with a zero shift
        if lay.notDoneOnce:
            lay.notDoneOnce = False
            lay.ratio = np.zeros(2)
            lay.ratio[0] = lay.vX * self.telescope.samplingTime / lay.pixel_size
            lay.ratio[1] = lay.vY * self.telescope.samplingTime / lay.pixel_size
            lay.buff = np.zeros(2)
        ratio = lay.ratio if shift is None else np.asarray(shift, np.float64)
        tmp = np.abs(ratio)
        tmp[np.isinf(tmp)] = 0
        nScreens = tmp.astype('int')
        step = np.zeros(2)
        for _ in range(nScreens.min()):
            step[0] = 1
            step[1] = 1
            step = step * np.sign(ratio)
            self._add_row(lay, step)
        for _ in range(nScreens.max() - nScreens.min()):
            step[0] = 1
            step[1] = 1
            step = step * np.sign(ratio)
            step[np.where(nScreens == nScreens.min())] = 0
            self._add_row(lay, step)
        sub = np.zeros(2)
        sub[0] = (np.abs(ratio[0]) % 1) * np.sign(ratio[0])
        sub[1] = (np.abs(ratio[1]) % 1) * np.sign(ratio[1])
        lay.buff += sub
        if np.abs(lay.buff[0]) >= 1 or np.abs(lay.buff[1]) >= 1:
            step[0] = 1 * np.sign(lay.buff[0])
            step[1] = 1 * np.sign(lay.buff[1])
            step[np.where(np.abs(lay.buff) < 1)] = 0
            self._add_row(lay, step)
        lay.buff[0] = (np.abs(lay.buff[0]) % 1) * np.sign(lay.buff[0])
        lay.buff[1] = (np.abs(lay.buff[1]) % 1) * np.sign(lay.buff[1])
        lay.shift = (float(lay.buff[0]), float(lay.buff[1]))

    def _compute_opd(self):
        """Sum of the shifted, clipped, footprint-cropped layers in metres (Atmosphere.py:513-516, 589-622)."""
        n = self.telescope.resolution
        if self._OPD_dev is None or self._OPD_dev.shape[0] != n:
            self._OPD_dev = torch.empty((n, n), dtype=torch.float32, device=self.device)
        for i in range(self.nLayer):
            lay = getattr(self, 'layer_' + str(i + 1))
            r0 = lay.center - n // 2 + 1
            w = self.wavelength / 2 / np.pi * np.sqrt(self.fractionalR0[i])
            tx, ty = getattr(lay, 'shift', (0.0, 0.0))
            _lib.call('oopao_atm_layer_opd', ptr(lay.map), lay.M, ptr(lay.mm), ctypes.c_double(tx),
                      ctypes.c_double(ty), r0, r0, ctypes.c_double(w), ptr(self._OPD_dev), n, 1 if i else 0,
                      stream_ptr(self.device))

    def update(self, OPD=None):
        """Atmosphere.update (Atmosphere.py:518-531)."""
        if self.hasNotBeenInitialized:
            raise OopaoError('The Atmosphere object needs to be initialised using the initialiseAtmosphere()')
        if OPD is None:
            self.is_user_defined_opd = False
            for i in range(self.nLayer):
                self.updateLayer(getattr(self, 'layer_' + str(i + 1)))
            self._compute_opd()
        else:
            self.is_user_defined_opd = True
            self.user_defined_opd = OPD
        if self.telescope.isPaired:
            self.telescope.src ** self * self.telescope

    # ------------------------------------------------------------------ relay
    def relay(self, src):
        """Atmosphere.relay (Atmosphere.py:533-581), geometric propagation of an on-axis source."""
        self.src = src
        self.src_list = [src] if src.tag == 'source' else src.src
        if self.is_user_defined_opd:
            warning('User-Defined OPD are only propagated once in the Atmosphere class.')
            opd = torch.as_tensor(np.asarray(self.user_defined_opd), dtype=torch.float32, device=self.device)
            self.is_user_defined_opd = False
        else:
            opd = self._OPD_dev
        n = self.telescope.resolution
        user = opd is not self._OPD_dev
        self._layer_full = {}
        self._src_opd = []
        for s in self.src_list:
            if s.coordinates[0] > getattr(self, 'fov', 0) / 2:
                raise OopaoError(f'Source zenith ({s.coordinates[0]}") outside of fov!')
            s.through_atm = True
            s.optical_path.append([self.tag, self])
            s._scintillation_no_pupil = torch.ones((n, n), dtype=torch.float32, device=self.device)
            m = s._mask_dev()
            s._scintillation = s._scintillation_no_pupil if m is None else s._scintillation_no_pupil * m
            s._refresh_intensity()
            if self.angular_spectrum_propagation and not user:
                o, inten = self._scintillation(s)
                s._scintillation_no_pupil = inten
                s._scintillation = inten if m is None else inten * m
                s._refresh_intensity()
            else:
                o = opd.clone() if (user or self._on_axis(s)) else self._source_opd(s)
            self._src_opd.append(o)
            s._set_OPD_no_pupil_dev(o)
        if self.angular_spectrum_propagation and not user:
            self.scintillation_map = np.squeeze(np.array([to_host(x._scintillation_no_pupil) for x in self.src_list]))

    # ------------------------------------------------------- off-axis / LGS
    @staticmethod
    def _on_axis(s):
        return s.coordinates[0] == 0 and not np.isfinite(s.altitude) and s.chromatic_shift is None

    def _full_layer(self, i):
        """layer.OPD of the reference (the shifted, clipped screen, [res, res] radians at 500 nm), once per relay."""
        f = self._layer_full.get(i)
        if f is None:
            lay = getattr(self, 'layer_' + str(i + 1))
            f = torch.empty((lay.resolution, lay.resolution), dtype=torch.float32, device=self.device)
            tx, ty = getattr(lay, 'shift', (0.0, 0.0))
            _lib.call('oopao_atm_layer_opd', ptr(lay.map), lay.M, ptr(lay.mm), ctypes.c_double(tx),
                      ctypes.c_double(ty), 1, 1, ctypes.c_double(1.0), ptr(f), lay.resolution, 0,
                      stream_ptr(self.device))
            self._layer_full[i] = f
        return f

    def footprint(self, lay, s, i_layer):
        """(row origin, column origin, sub-pixel shift x, y) of a source's footprint on a layer
        (set_pupil_footprint, Atmosphere.py:423-445)."""
        if s.chromatic_shift is not None:
            if len(s.chromatic_shift) != self.nLayer:
                raise OopaoError('The chromatic_shift property is expected to be the same length as the number of '
                                 'atmospheric layer. ')
            ch = s.chromatic_shift[i_layer]
        else:
            ch = 0
        rho = lay.altitude * np.tan((s.coordinates[0] + ch) / self.rad2arcsec) * lay.resolution / lay.D
        phi = np.deg2rad(s.coordinates[1])
        x_z, y_z = rho * np.cos(phi), rho * np.sin(phi)
        n = self.telescope.resolution
        cx = int(y_z) + lay.resolution // 2
        cy = int(x_z) + lay.resolution // 2
        return cx - n // 2, cy - n // 2, int(x_z) - x_z, int(y_z) - y_z

    def _source_layer_opd(self, s, i, out, accumulate):
        """Layer i's OPD (metres) on the footprint of source s into out [n, n] (fill_OPD_support,
        Atmosphere.py:589-622), added to out when ``accumulate``."""
        n = self.telescope.resolution
        st = stream_ptr(self.device)
        lay = getattr(self, 'layer_' + str(i + 1))
        if s.altitude <= lay.altitude:
            raise OopaoError('The source altitude (' + str(s.altitude) + ' m) is below or at the same altitude as '
                             'the atmosphere layer (' + str(lay.altitude) + ' m)')
        w = self.wavelength / 2 / np.pi * np.sqrt(self.fractionalR0[i])
        if self._on_axis(s):
            tx, ty = getattr(lay, 'shift', (0.0, 0.0))
            r0 = lay.center - n // 2 + 1
            _lib.call('oopao_atm_layer_opd', ptr(lay.map), lay.M, ptr(lay.mm), ctypes.c_double(tx),
                      ctypes.c_double(ty), r0, r0, ctypes.c_double(w), ptr(out), n, int(accumulate), st)
            return
        full = self._full_layer(i)
        r0, c0, esx, esy = self.footprint(lay, s, i)
        res = lay.resolution
        if np.isfinite(s.altitude):
            # footprint (translated by the sub-pixel remainder), then the cone magnification
            # m = (H - h) / H about the footprint centre, with the reference's half-pixel alignment
            sub = torch.empty((n, n), dtype=torch.float32, device=self.device)
            _lib.call('oopao_bilinear_affine', ptr(full), res, res, ptr(sub), n, n, ctypes.c_double(1.0),
                      ctypes.c_double(r0 - esy), ctypes.c_double(c0 - esx), ctypes.c_double(1.0), 0, st)
            m = (s.altitude - lay.altitude) / s.altitude
            ratio = 1.0 / m
            e = (ratio % 1) / 2 + (np.floor(ratio) - 1) * 0.5
            b = -(n / 2 + e) / ratio + n / 2
            _lib.call('oopao_bilinear_affine', ptr(sub), n, n, ptr(out), n, n, ctypes.c_double(1.0 / ratio),
                      ctypes.c_double(b), ctypes.c_double(b), ctypes.c_double(w), int(accumulate), st)
        else:
            _lib.call('oopao_bilinear_affine', ptr(full), res, res, ptr(out), n, n, ctypes.c_double(1.0),
                      ctypes.c_double(r0 - esy), ctypes.c_double(c0 - esx), ctypes.c_double(w), int(accumulate), st)

    def _source_opd(self, s):
        """Sum over layers of the source's footprint (fill_OPD_support, Atmosphere.py:589-622), metres."""
        n = self.telescope.resolution
        out = torch.empty((n, n), dtype=torch.float32, device=self.device)
        for i in range(self.nLayer):
            self._source_layer_opd(s, i, out, i > 0)
        return out

    # ---------------------------------------------------- scintillation (ASM)
    def _transfer_function(self, wvl, distance):
        """H = exp(-i pi lambda z (fx^2 + fy^2)) on the centred frequency grid (Atmosphere.py:938-972)."""
        n = self.telescope.resolution
        pitch = getattr(self.telescope, 'pixelSize', self.telescope.D / n)
        key = (n, float(pitch), float(wvl), float(distance))
        H = self._transfer.get(key)
        if H is None:
            f = np.arange(-n / 2, n / 2) / (n * pitch)
            fx, fy = np.meshgrid(f, f)
            Hh = np.exp(-1j * np.pi * wvl * distance * (fx ** 2 + fy ** 2))
            H = torch.from_numpy(Hh.astype(np.complex64)).to(self.device)
            _lib.prepare_tables(n, self.device)
            self._transfer[key] = H
        return H

    def _scintillation(self, s):
        """Field of source s through the layers, highest first, each multiplied in and propagated to the
        next one (or the ground) by the angular spectrum (fill_scintillation_support,
        Atmosphere.py:644-662), then (OPD, intensity) as set_scintillation_support (664-706)."""
        self.check_fresnel_sampling()
        n = self.telescope.resolution
        dev = self.device
        st = stream_ptr(dev)
        k = 2 * np.pi / s.wavelength
        order = sorted(range(self.nLayer), key=lambda i: getattr(self, 'layer_' + str(i + 1)).altitude, reverse=True)
        field = torch.empty((n, n), dtype=torch.complex64, device=dev)
        lopd = torch.empty((n, n), dtype=torch.float32, device=dev)
        nb = ctypes.c_size_t(0)
        _lib.call('oopao_asm_bytes', n, ctypes.byref(nb))
        ws = torch.empty(nb.value, dtype=torch.uint8, device=dev)
        alts = [getattr(self, 'layer_' + str(i + 1)).altitude for i in order]
        for j, i in enumerate(order):
            self._source_layer_opd(s, i, lopd, False)
            _lib.call('oopao_field_phase', ptr(lopd), ctypes.c_float(k), ptr(field), n * n, int(j == 0), st)
            dist = alts[j] - alts[j + 1] if j < len(order) - 1 else alts[j]
            if dist > 1e-6:
                _lib.call('oopao_asm_propagate', ptr(field), ptr(self._transfer_function(s.wavelength, dist)), n,
                          ptr(ws), ws.numel(), st)
        inten = torch.empty((n, n), dtype=torch.float32, device=dev)
        if self.geometric_phase_backup:
            geo = self._source_opd(s)
            _lib.call('oopao_field_readout', ptr(field), None, ctypes.c_float(k), ptr(inten), None, n * n, st)
            return geo, inten
        out = torch.empty((n, n), dtype=torch.float32, device=dev)
        geo = self._source_opd(s) if self.unwrap_diffractive_phase else None
        _lib.call('oopao_field_readout', ptr(field), None if geo is None else ptr(geo), ctypes.c_float(k), ptr(inten),
                  ptr(out), n * n, st)
        if geo is not None:
            pupil = self.telescope._pupil_dev > 0
            delta = (out - geo)[pupil] * k
            if delta.numel() and float(delta.max() - delta.min()) > 1.9 * np.pi:
                warning('Diffractive residual wrapped inside the pupil! Scintillation is too strong for perfect '
                        'unwrapping.')
        return out, inten

    def update_rytov_variance(self):
        """Rytov variance of the profile at rytov_wvl (Atmosphere.py:917-936)."""
        if not hasattr(self, 'nLayer') or not hasattr(self, 'fractionalR0') or not hasattr(self, 'altitude'):
            return
        k_t = 2 * np.pi / self.rytov_wvl
        r0_t = self.r0 * (self.rytov_wvl / self.wavelength) ** (6 / 5)
        cn2_total = (r0_t ** (-5 / 3)) / (0.423 * k_t ** 2)
        cur = 0.0
        for i in range(self.nLayer):
            if self.altitude[i] > 1e-3:
                cur += (cn2_total * self.fractionalR0[i]) * (self.altitude[i] ** (5 / 6))
        self._rytov_var = cur * 2.2524 * (k_t ** (7 / 6))
        if self.angular_spectrum_propagation:
            if self._rytov_var > 0.3 and not self._saturation_warned:
                warning(f"Atmospheric conditions degraded. Rytov variance ({self._rytov_var:.2f}) exceeds 0.3 limit! "
                        "Entering saturation regime.")
                self._saturation_warned = True
            elif self._rytov_var <= 0.3:
                self._saturation_warned = False

    @property
    def rytov_var(self):
        if not self.angular_spectrum_propagation:
            return 0.0
        if not hasattr(self, '_rytov_var'):
            self.update_rytov_variance()
        return self._rytov_var

    @rytov_var.setter
    def rytov_var(self, val):
        """Scale r0 (and the current layer screens) to reach a target Rytov variance (Atmosphere.py:1172-1205)."""
        if not self.angular_spectrum_propagation:
            warning("Cannot set Rytov variance because atm.angular_spectrum_propagation is False.")
            return
        val = float(val)
        if val <= 0:
            return
        if val > 0.3:
            warning(f"Target Rytov variance ({val:.2f}) exceeds the weak fluctuation limit (0.3). The simulation is "
                    "entering the moderate/strong scintillation regime (saturation). Phase unwrapping and scaling "
                    "behaviors may become unstable.")
        self.update_rytov_variance()
        cur = self._rytov_var
        if cur > 0:
            ratio = val / cur
            print(f"--- Scaling Rytov variance to {val:.4f} (Ratio: {ratio:.3f}) ---")
            self.r0 = self.r0 * (ratio ** (-3 / 5))
            scale = float(np.sqrt(ratio))
            if not self.hasNotBeenInitialized:
                for i in range(self.nLayer):
                    lay = getattr(self, 'layer_' + str(i + 1))
                    lay.map.mul_(scale)
                    lay.mm[:2].mul_(scale)
                self._compute_opd()
            self.update_rytov_variance()
        else:
            warning("Cannot scale the Rytov variance: every layer is at the ground.")

    def check_fresnel_sampling(self):
        """Grid checks for the Fresnel propagation (Atmosphere.py:974-1010): warnings only."""
        if not self.angular_spectrum_propagation or self.sampling_checked:
            return
        tel = self.telescope
        D = getattr(tel, 'initial_D', tel.D)
        N = tel.resolution
        delta = getattr(tel, 'pixelSize', tel.D / N)
        alts = sorted(list(self.altitude) + [0.0], reverse=True)
        max_step = max([alts[i] - alts[i + 1] for i in range(len(alts) - 1)]) if len(alts) > 1 else 0
        z_max = max(self.altitude) if len(self.altitude) else 0
        for src in self.src_list:
            wvl = src.wavelength
            r0_w = self._r0 * (wvl / self.wavelength) ** (6 / 5)
            D_turb = 4.0 * (wvl * z_max) / r0_w
            D_total = D + 2.0 * D_turb
            delta_req = min(r0_w / 4.0, (wvl * z_max) / D_total) if z_max > 0 else r0_w / 4.0
            N_min = (D_total / delta) + (wvl * z_max) / (delta ** 2)
            z_asm = (N * delta ** 2) / wvl
            if delta > delta_req:
                warning(f"Fresnel [wvl={wvl*1e9:.0f}nm] - Pixel scale ({delta*1000:.1f} mm) too large! "
                        f"Limit is {delta_req*1000:.1f} mm.")
            if N < N_min:
                warning(f"Fresnel Aliasing [wvl={wvl*1e9:.0f}nm] - Grid N={N} too small (Physics requires "
                        f"N >= {int(np.ceil(N_min))}).")
            if max_step > z_asm:
                warning(f"Fresnel ASM Limit [wvl={wvl*1e9:.0f}nm] - Distance between layers ({max_step/1000:.1f} km) "
                        f"> limit ({z_asm/1000:.1f} km).")
        self.sampling_checked = True

    # -------------------------------------------------------------- properties
    @property
    def OPD(self):
        """On-axis OPD, or the per-source OPDs [nSrc, n, n] after relaying an asterism (Atmosphere.py:624-629)."""
        lst = getattr(self, '_src_opd', None)
        if lst and (len(lst) > 1 or not self._on_axis(self.src_list[0])):
            return np.squeeze(np.array([to_host(o) for o in lst]))
        return to_host(self._OPD_dev)

    @property
    def OPD_no_pupil(self):
        return self.OPD

    def layer_opd(self, i_layer):
        """Full shifted screen of one layer in radians at 500 nm ([res, res], layer.OPD of the reference)."""
        lay = getattr(self, 'layer_' + str(i_layer + 1))
        out = torch.empty((lay.resolution, lay.resolution), dtype=torch.float32, device=self.device)
        tx, ty = getattr(lay, 'shift', (0.0, 0.0))
        _lib.call('oopao_atm_layer_opd', ptr(lay.map), lay.M, ptr(lay.mm), ctypes.c_double(tx), ctypes.c_double(ty),
                  1, 1, ctypes.c_double(1.0), ptr(out), lay.resolution, 0, stream_ptr(self.device))
        return out.double().cpu().numpy()

    @property
    def r0(self):
        return self._r0

    @r0.setter
    def r0(self, val):
        """r0 setter (Atmosphere.py:1017-1036): rescales the covariances and recomputes B."""
        self.sampling_checked = False
        self._r0 = val
        self.r0_zenith = val * np.sin(np.radians(self.elevation)) ** (-3 / 5)
        if self.hasNotBeenInitialized is False:
            print('Updating the Atmosphere covariance matrices...')
            self.seeingArcsec = self.rad2arcsec * (self.wavelength / val)
            self.cn2 = (self.r0 ** (-5. / 3) / (0.423 * (2 * np.pi / self.wavelength) ** 2)) / np.max([1, np.max(self.altitude)])
            for i in range(self.nLayer):
                lay = getattr(self, 'layer_' + str(i + 1))
                lay.B = extrusion_operators(lay.ZZt, lay.ZXt, lay.XXt, lay.ZZt_inv, self.r0_def, self.r0)[1]
                self._set_B(lay)

    @property
    def windSpeed(self):
        return self._windSpeed

    @windSpeed.setter
    def windSpeed(self, val):
        self._windSpeed = val
        if self.hasNotBeenInitialized is False:
            if len(val) != self.nLayer:
                raise OopaoError('Wrong value for the wind-speed! Make sure that you inpute a wind-speed for each layer')
            # the setter's own V0 expression (Atmosphere.py:1067), not the constructor's (197)
            self.V0 = (np.sum(np.asarray(self.fractionalR0) * np.asarray(self.windSpeed)) ** (5 / 3)) ** (3 / 5)
            self.tau0 = 0.31 * self.r0 / self.V0
            for i in range(self.nLayer):
                lay = getattr(self, 'layer_' + str(i + 1))
                lay.windSpeed = val[i]
                self._update_velocity(lay)

    @property
    def windDirection(self):
        return self._windDirection

    @windDirection.setter
    def windDirection(self, val):
        self._windDirection = val
        if self.hasNotBeenInitialized is False:
            if len(val) != self.nLayer:
                raise OopaoError('Wrong value for the wind-speed! Make sure that you inpute a wind-direction for each layer')
            for i in range(self.nLayer):
                lay = getattr(self, 'layer_' + str(i + 1))
                lay.direction = val[i]
                self._update_velocity(lay)

    @property
    def L0(self):
        return self._L0

    @L0.setter
    def L0(self, val):
        """L0 setter (Atmosphere.py:1042-1052): rebuilds the layers (covariances, screens) for the new
        outer scale."""
        self._L0 = val
        if self.hasNotBeenInitialized is False:
            print('Updating the Atmosphere covariance matrices...')
            self.hasNotBeenInitialized = True
            self.initializeAtmosphere(self.telescope)

    @property
    def fractionalR0(self):
        return self._fractionalR0

    @fractionalR0.setter
    def fractionalR0(self, val):
        """fractionalR0 setter (Atmosphere.py:1122-1133): V0 and tau0 (the layer weights of the OPD sum
        read this list), then the Rytov variance."""
        self._fractionalR0 = val
        if self.hasNotBeenInitialized is False:
            if len(val) != self.nLayer:
                raise OopaoError('Wrong value for the fractional r0 ! Make sure that you inpute a fractional r0 for '
                                 'each layer! If you want to change the number of layer, re-generate a new atmosphere '
                                 'object.')
            print('Updating the fractional R0...BEWARE COMPLETE THE RECOMPUTATION...NOT ONLY V0 and Tau0 !')
            self.V0 = (np.sum(np.asarray(self.fractionalR0) * np.asarray(self.windSpeed)) ** (5 / 3)) ** (3 / 5)
            self.tau0 = 0.31 * self.r0 / self.V0
        self.update_rytov_variance()

    @property
    def elevation(self):
        return self._elevation

    @elevation.setter
    def elevation(self, val):
        """elevation setter (Atmosphere.py:1139-1161): clamps to 5 deg, moves the layers to the slant
        distances altitude_zenith / sin(el), scales r0 from its zenith value and rebuilds the layers."""
        self.sampling_checked = True
        val = float(val)
        if val < 5.0:
            warning("Very low elevation (< 5 deg). Clamping to 5 deg.")
            val = 5.0
        if hasattr(self, '_elevation') and val == self._elevation:
            return
        self._elevation = val
        el = np.radians(val)
        print(f"--- Atmosphere Configuration : Elev={val:.1f}\u00b0 ---")
        self.altitude = [h_z / np.sin(el) for h_z in self.altitude_zenith]
        self._r0 = self.r0_zenith * np.sin(el) ** (3 / 5)
        print(f" -> New r0 : {self._r0:.4f} m (Zenith reference: {self.r0_zenith:.4f} m)")
        if self.hasNotBeenInitialized is False:
            self.hasNotBeenInitialized = True
            if getattr(self, 'telescope', None) is None:
                raise OopaoError("Warning: Atmosphere not yet linked to a telescope. Call initializeAtmosphere() "
                                 "manually.")
            self.initializeAtmosphere(self.telescope, compute_covariance=getattr(self, 'compute_covariance', True))
        self.update_rytov_variance()

    def _update_velocity(self, lay):
        lay.vY = lay.windSpeed * np.cos(np.deg2rad(lay.direction))
        lay.vX = lay.windSpeed * np.sin(np.deg2rad(lay.direction))
        lay.ratio[0] = lay.vX * self.telescope.samplingTime / lay.pixel_size
        lay.ratio[1] = lay.vY * self.telescope.samplingTime / lay.pixel_size

    # --------------------------------------------------------------- operators
    def __mul__(self, obj):
        if obj.tag in ('telescope', 'source', 'asterism'):
            if obj.tag == 'telescope':
                self.telescope = obj
                self.relay(obj.src)
            else:
                self.relay(obj)
            return obj
        raise OopaoError('The Atmosphere can be multiplied only with a Telescope or a Source object!')

    def print_properties(self):
        print(self)

    def __repr__(self):
        s = '\n------------ Atmosphere ------------\n'
        s += f"{'r0 @500nm [m]':<25s}|{self.r0:^9.3f}\n{'L0 [m]':<25s}|{self.L0:^9.1f}\n"
        s += f"{'Layers':<25s}|{self.nLayer:^9d}\n"
        for i, t_b in enumerate(self.t_boiling):
            evo = 'Frozen Flow' if t_b is None else f'Boiling: t = {t_b:.4g} s'
            s += f"{'Layer %d' % (i + 1):<25s}|{evo:^24s}\n"
        return s
