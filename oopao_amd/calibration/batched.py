"""Push-pull interaction matrices of several PWFS systems in one HIP graph, with an autotuner.

The reference calibrates one system at a time (calibration/InteractionMatrix.py:13-169): for
each batch of modes it sets dm.coefs = M2C[:, batch] * stroke, propagates ngs**tel*dm*wfs,
repeats with -stroke, and stores 0.5 (s_push - s_pull) / stroke as the columns. Here every
batch of every system (different pupil sampling, modulation points and DMs) is recorded once
into a single HIP graph; one replay then runs the whole calibration with no host work between
the kernels:

    per system, per batch of B modes:
        OPD of the B push commands (separable or dense DM route), OPD of the pulls = -push
        times the pupil, PWFS row pass, column pass, slopesMaps of the 2 B frames,
        D[batch] = 0.5 (s_push - s_pull) / stroke.

``PWFSAutotuner`` picks, per PWFS geometry, the FFT plan of each pass (lanes per transform, which
fixes the radix plan, the values per lane and the occupancy; ``PWFSEngine.plans()`` lists the compiled
ones), the column-pass chunk (modulation points per chunk), the row-pass task size (tip tables per
task) and the batch size by timing candidate settings on the device; choices are cached per geometry
for the process and can be saved to / loaded from a JSON file.
"""
import json
import math

import numpy as np
import torch

from ..tools import OopaoError


def _event_ms(fn, reps):
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


class PWFSAutotuner:
    """Device timing of PWFS tile settings; results cached per (device, n, N, T, tip tables)."""

    def __init__(self, batch_candidates=(4, 8, 16, 32, 64), memory_fraction=0.25):
        self.batch_candidates = tuple(batch_candidates)
        self.memory_fraction = memory_fraction
        self.cache = {}

    @staticmethod
    def key(wfs):
        e = wfs.engine
        return '%s|n=%d|N=%d|T=%d|P=%d' % (torch.cuda.get_device_name(wfs.device), e.n, e.N, e.T, e.P)

    def save(self, path):
        with open(path, 'w') as f:
            json.dump(self.cache, f, indent=1, sort_keys=True)

    def load(self, path):
        with open(path) as f:
            self.cache.update(json.load(f))

    def _time_pass(self, wfs, opd, amp, scale, B, which, reps=3):
        """ms of one pass ('rows' or 'cols') for B screens with the engine's current settings."""
        e = wfs.engine
        buf = e.buffers(B)
        frames = torch.empty((B, e.cam, e.cam), dtype=torch.float32, device=wfs.device)
        bands = wfs.slope_bands()
        e.rows(opd[:B], scale, amp, buf, bands)
        if which == 'rows':
            return _event_ms(lambda: e.rows(opd[:B], scale, amp, buf, bands), reps)
        return _event_ms(lambda: e.cols(buf, frames, bands), reps)

    def _time(self, wfs, opd, amp, scale, B, reps=3):
        e, se = wfs.engine, wfs.slopes_engine
        buf = e.buffers(B)
        frames = torch.empty((B, e.cam, e.cam), dtype=torch.float32, device=wfs.device)
        s2d = torch.empty((B, 2 * se.w, se.w), dtype=torch.float32, device=wfs.device)
        s = torch.empty((B, se.nsig), dtype=torch.float32, device=wfs.device)

        bands = wfs.slope_bands()

        def run():
            e.rows(opd[:B], scale, amp, buf, bands)
            e.cols(buf, frames, bands)
            se.slopes(frames, s2d, s)
        return _event_ms(run, reps) / B

    def tune(self, wfs, amp, scale, opd_sample, max_batch=None):
        """Best {'nchunk', 'pts_per_task', 'batch', 'ms_per_frame'} for this sensor (applied to its engine)."""
        e = wfs.engine
        if e.mode != 'fast':
            raise OopaoError('the autotuner covers the separable PWFS path (N = 4n geometries)')
        k = self.key(wfs)
        if k in self.cache:
            best = self.cache[k]
            e.nchunk_user, e.pts_per_task_user = best['nchunk'], best['pts_per_task']
            e.set_plans(best.get('lanes_rows', 0), best.get('lanes_cols', 0))
            return best
        free = torch.cuda.mem_get_info(wfs.device)[0]
        per = max(1, e.bytes_per_screen())
        bmax = max(1, int(self.memory_fraction * free // per))
        if max_batch:
            bmax = min(bmax, max_batch)
        batches = [b for b in self.batch_candidates if b <= bmax] or [1]
        B0 = min(8, batches[-1])
        opd = opd_sample.expand(max(batches), -1, -1).contiguous()
        # FFT plan of each pass at a fixed batch and the default tiles (each plan changes only its pass)
        rows_plans, cols_plans = e.plans()
        e.nchunk_user = e.pts_per_task_user = None
        plan_ms = {}
        for which, plans in (('rows', rows_plans), ('cols', cols_plans)):
            for lanes in plans:
                e.set_plans(lanes if which == 'rows' else e.lanes_rows, lanes if which == 'cols' else e.lanes_cols)
                plan_ms[(which, lanes)] = self._time_pass(wfs, opd, amp, scale, B0, which)
            pick = min(plans, key=lambda v: plan_ms[(which, v)])
            e.set_plans(pick if which == 'rows' else e.lanes_rows, pick if which == 'cols' else e.lanes_cols)
        # tile sizes at a fixed batch: modulation points per column-pass chunk, the cap of tip tables per
        # rows-pass workgroup
        chunks = sorted({max(1, math.ceil(e.T / c)) for c in (4, 6, 9, 12)})
        tasks = sorted({max(1, min(e.P, p)) for p in (2, 4, 8)})
        trials = []
        for nc in chunks:
            for pt in tasks:
                e.nchunk_user, e.pts_per_task_user = nc, pt
                trials.append((self._time(wfs, opd, amp, scale, B0), nc, pt))
        t0, nc, pt = min(trials)
        e.nchunk_user, e.pts_per_task_user = nc, pt
        # batch size with the chosen tiles
        bt = [(self._time(wfs, opd, amp, scale, b), b) for b in batches]
        tb, b = min(bt)
        best = {'nchunk': nc, 'pts_per_task': pt, 'batch': b, 'ms_per_frame': round(tb, 5),
                'lanes_rows': e.lanes_rows, 'lanes_cols': e.lanes_cols,
                'plan_ms': {'%s_%d' % key: round(v, 5) for key, v in plan_ms.items()},
                'trials': len(trials) + len(bt) + len(plan_ms)}
        self.cache[k] = best
        return best


class CalibrationSystem:
    """One (source, telescope, DM, PWFS, M2C) system of a batched push-pull calibration."""

    def __init__(self, ngs, tel, dm, wfs, M2C, stroke):
        if wfs.postProcessing != 'slopesMaps' or wfs.binning != 1:
            raise OopaoError('batched calibration covers slopesMaps sensors without binning')
        if wfs.engine.mode != 'fast':
            raise OopaoError('batched calibration covers the separable PWFS path (N = 4n geometries)')
        self.ngs, self.tel, self.dm, self.wfs = ngs, tel, dm, wfs
        self.stroke = float(stroke)
        self.dev = wfs.device
        M = M2C if isinstance(M2C, torch.Tensor) else torch.from_numpy(np.asarray(M2C, dtype=np.float64))
        if M.dim() == 1:
            M = M[:, None]
        if M.shape[0] != dm.nValidAct:
            raise OopaoError('M2C has %d rows, the DM has %d actuators' % (M.shape[0], dm.nValidAct))
        self.cmd = (M.to(self.dev, torch.float64) * self.stroke).to(torch.float32).contiguous()
        self.nModes = self.cmd.shape[1]
        tel.isPaired = False
        ngs ** tel
        self.amp = wfs._amp()
        self.scale = 2 * np.pi / ngs.wavelength
        self.pupil = tel._pupil_dev.float().contiguous()
        self.batch = None
        self.D = torch.zeros((self.nModes, wfs.nSignal), dtype=torch.float32, device=self.dev)  # rows = modes

    def tune(self, tuner):
        flat = torch.zeros((1, self.tel.resolution, self.tel.resolution), dtype=torch.float32, device=self.dev)
        rs = np.random.RandomState(0)
        flat[0] = torch.from_numpy((rs.randn(self.tel.resolution, self.tel.resolution) * 20e-9).astype(np.float32))
        best = tuner.tune(self.wfs, self.amp, self.scale, flat * self.pupil, max_batch=64)
        self.batch = max(1, best['batch'] // 2)     # a batch of B modes measures 2 B frames
        return best

    def prepare(self, batch=None):
        """Allocate every buffer the recorded work touches (nothing is allocated while recording)."""
        B = batch or self.batch or 8
        self.batch = B
        e, se = self.wfs.engine, self.wfs.slopes_engine
        n = self.tel.resolution
        self.batches = [(i0, min(self.nModes, i0 + B)) for i0 in range(0, self.nModes, B)]
        self.opd2 = torch.empty((2 * B, n, n), dtype=torch.float32, device=self.dev)
        self.frames = torch.empty((2 * B, e.cam, e.cam), dtype=torch.float32, device=self.dev)
        self.s2d = torch.empty((2 * B, 2 * se.w, se.w), dtype=torch.float32, device=self.dev)
        self.s = torch.empty((2 * B, se.nsig), dtype=torch.float32, device=self.dev)
        self.bufs = {}
        for i0, i1 in self.batches:
            b2 = 2 * (i1 - i0)
            if b2 not in self.bufs:
                self.bufs[b2] = e.buffers(b2)
                se.slopes(self.frames[:b2], self.s2d[:b2], self.s[:b2])   # grows the slopes workspace now
        self.neg_pupil = (-self.pupil).contiguous()
        self.bands = self.wfs.slope_bands()     # only the slopes are kept: quadrant columns of the frames
        self.half_over_stroke = 0.5 / self.stroke

    def record(self):
        """The work of every batch, on the current stream (recorded into the graph by the caller)."""
        e, se = self.wfs.engine, self.wfs.slopes_engine
        for i0, i1 in self.batches:
            b = i1 - i0
            opd = self.dm.opd_from_coefs(self.cmd[:, i0:i1])
            opd = opd if opd.dim() == 3 else opd[None]
            torch.mul(opd, self.pupil, out=self.opd2[:b])
            torch.mul(opd, self.neg_pupil, out=self.opd2[b:2 * b])
            buf = self.bufs[2 * b]
            e.rows(self.opd2[:2 * b], self.scale, self.amp, buf, self.bands)
            e.cols(buf, self.frames[:2 * b], self.bands)
            se.slopes(self.frames[:2 * b], self.s2d[:2 * b], self.s[:2 * b])
            torch.sub(self.s[:b], self.s[b:2 * b], out=self.D[i0:i1])
            self.D[i0:i1].mul_(self.half_over_stroke)

    @property
    def interaction_matrix(self):
        """D [nSignal, nModes] float32 device tensor (columns = 0.5 (s_push - s_pull) / stroke)."""
        return self.D.t()


class BatchedCalibration:
    """Push-pull calibration of several systems recorded into one HIP graph.

    ``systems``: list of (ngs, tel, dm, wfs, M2C). ``tune=True`` runs the autotuner first;
    ``batch`` forces the modes per batch instead.
    """

    def __init__(self, systems, stroke, tune=True, batch=None, tuner=None):
        if not systems:
            raise OopaoError('no calibration system given')
        self.systems = [CalibrationSystem(*s, stroke=stroke) for s in systems]
        self.tuner = tuner or PWFSAutotuner()
        self.tuning = []
        for cs in self.systems:
            if tune and batch is None:
                self.tuning.append(cs.tune(self.tuner))
            cs.prepare(batch)
        self.graph = None

    def run_eager(self):
        for cs in self.systems:
            cs.record()
        return [cs.interaction_matrix for cs in self.systems]

    def capture(self, concurrent=True):
        """Warm up on a side stream, then record every batch of every system into one graph. With
        ``concurrent`` each system is recorded on its own stream, forked from the capture stream and
        joined back into it, so the graph holds one independent branch per system and the smaller
        sensors' kernels run beside the larger ones' (the systems share no buffers; every stream has the
        default priority)."""
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self.run_eager()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        if concurrent and len(self.systems) > 1:
            branches = [torch.cuda.Stream() for _ in self.systems]
        with torch.cuda.graph(self.graph):
            if concurrent and len(self.systems) > 1:
                cap = torch.cuda.current_stream()
                for st, cs in zip(branches, self.systems):
                    st.wait_stream(cap)
                    with torch.cuda.stream(st):
                        cs.record()
                for st in branches:
                    cap.wait_stream(st)
            else:
                for cs in self.systems:
                    cs.record()
        return self.graph

    def run(self):
        """One graph replay: the interaction matrices [nSignal, nModes] of all systems."""
        if self.graph is None:
            self.capture()
        self.graph.replay()
        return [cs.interaction_matrix for cs in self.systems]

    def vaults(self, nTrunc=0):
        """CalibrationVault (float64 SVD pseudo-inverse on the GPU) of each interaction matrix."""
        from .CalibrationVault import CalibrationVault
        out = []
        for cs in self.systems:
            cv = CalibrationVault(cs.interaction_matrix.double(), nTrunc=nTrunc)
            out.append(cv)
        return out
