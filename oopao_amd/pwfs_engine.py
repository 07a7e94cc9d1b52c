"""Device tables and calls for the modulated Pyramid WFS (HIP library front end).

The reference builds, per modulation point i, the field
amp * exp(i (phi + TT_i)) on the n x n pupil, pads it to N x N, multiplies by
the PSF-centring phasor, and runs fft2 -> * mask -> ifft2 -> |.|^2
(Pyramid.py:612-646). This module factors the separable parts on the host in
float64 (rounded once to complex64):

* tx[i, x] = exp(i px_i Tip[x]) * phasor(c0 + x), ty[i, y] likewise with Tilt
  (Pyramid.py:334-344, 1129-1139);
* the mask phase m[k, l] = a[k] + b[l] (the default pyramid mask,
  Pyramid.py:453-480) gives mx[l] = exp(i b[l]) e^{-2 pi i l c0/N} / N and
  my[k] = exp(i a[k]) e^{-2 pi i k c0/N} / N, folding the zero-padding offset c0
  and the 1/N of each inverse transform.

Sensors outside that fast path (a non-separable mask: rooftop, rotation,
per-quadrant shifts, spatial filter; a padded length other than 4n; a binning
other than 4 x 4) run the general three-pass path of the library with the full
N x N complex64 mask (divided by N^2, stored column-major) and the pupil placed
at its offset in the padded support; tx and ty are the same tables.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._device import alloc_workspace
from .tools import OopaoError


def _stream_ptr(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def separable_mask_factors(m, tol=1e-9):
    """Split a mask phase m[k, l] = a[k] + b[l]; returns (a, b) or None if not separable."""
    a = m[:, 0].copy()
    b = m[0, :] - m[0, 0]
    resid = np.max(np.abs(m - a[:, None] - b[None, :]))
    scale = max(1.0, np.max(np.abs(m)))
    if resid > tol * scale:
        return None
    return a, b


def modulation_phasors(path, tip_lin, tilt_lin, N, n, psfCentering=True):
    """tx[T, n], ty[T, n] complex128 (see module docstring)."""
    c0 = N // 2 - n // 2
    X = np.arange(n) + c0
    if psfCentering:
        ph = np.exp(-1j * np.pi * (N + 1) / N * X)
    else:
        # fftshift of the spectrum = a phase ramp of N//2 cycles on the support ((-1)^X for even N)
        ph = np.exp(2j * np.pi * X * (N // 2) / N)
    tx = np.exp(1j * path[:, 0:1] * tip_lin[None, :]) * ph[None, :]
    ty = np.exp(1j * path[:, 1:2] * tilt_lin[None, :]) * ph[None, :]
    return tx, ty


def tip_groups(path, tol=1e-12):
    """Points sharing their tip value px (to ``tol`` relative): returns (order, pidx, reps) with
    ``order`` the points sorted so that each group is contiguous, ``pidx[k]`` the group of point
    order[k] and ``reps`` one member per group. On the default circular path points i and T - i
    share cos(theta) (Pyramid.py:1129-1132), so P = T/2 + 1."""
    px = np.asarray(path, np.float64)[:, 0]
    order = np.argsort(px, kind='stable')
    pidx = np.zeros(px.size, np.int32)
    reps = [int(order[0])]
    for k in range(1, px.size):
        a, b = px[order[k - 1]], px[order[k]]
        if b - a > tol * max(1.0, abs(a), abs(b)):
            reps.append(int(order[k]))
        pidx[k] = len(reps) - 1
    return order, pidx, np.asarray(reps)


def column_bands(bands, cam):
    """ctypes int[4] {a0, a1, b0, b1} for the C-ABI (None: every detector column). ``bands`` is a
    sequence of (start, stop) detector-column intervals; they are sorted and merged."""
    if bands is None:
        return None
    iv = sorted((int(a), int(b)) for a, b in bands if int(b) > int(a))
    merged = []
    for a, b in iv:
        if merged and a <= merged[-1][1]:
            merged[-1] = (merged[-1][0], max(merged[-1][1], b))
        else:
            merged.append((a, b))
    if not merged or len(merged) > 2 or merged[0][0] < 0 or merged[-1][1] > cam:
        raise OopaoError('column bands %s: need one or two intervals inside [0, %d)' % (list(bands), cam))
    if len(merged) == 1:
        merged.append((cam, cam))
    return (ctypes.c_int * 4)(merged[0][0], merged[0][1], merged[1][0], merged[1][1])


def factor_vectors(p, q, N, n):
    """my[N], mx[N] complex128 for a mask that factorises as p[k] q[l] (k: row, l: column), with the
    zero-padding offset shift and the 1/N of each inverse transform folded in (see module docstring)."""
    c0 = N // 2 - n // 2
    k = np.arange(N)
    shift = np.exp(-2j * np.pi * k * c0 / N) / N
    return np.asarray(p, np.complex128) * shift, np.asarray(q, np.complex128) * shift


def mask_vectors(a, b, N, n):
    """my[N], mx[N] complex128 from the separable mask phase (see module docstring)."""
    return factor_vectors(np.exp(1j * np.asarray(a)), np.exp(1j * np.asarray(b)), N, n)


def rank1_mask_factors(M, tol=1e-5):
    """(p, q) with M[k, l] = p[k] q[l] to ``tol`` of the largest entry when the complex mask M is of rank
    one, else None. A Bi-O Edge mask (a knife-edge amplitude along one axis times a tip / tilt phase,
    BioEdge.py:331-390) is; the factors are a column and a row of M through its largest entry."""
    M = np.asarray(M, np.complex128)
    k0, l0 = np.unravel_index(int(np.argmax(np.abs(M))), M.shape)
    piv = M[k0, l0]
    if piv == 0:
        return None
    p = M[:, l0] / piv
    q = M[k0, :]
    if np.max(np.abs(M - np.outer(p, q))) > tol * abs(piv):
        return None
    return p, q


class Workspace:
    """Device scratch that grows on demand. Engines whose calls run one after another on one stream
    (the four masks of a Bi-O Edge sensor) share one, so their peak memory is one workspace."""

    def __init__(self, device):
        self.device = device
        self.buf = None

    def get(self, nbytes):
        if self.buf is None or self.buf.numel() < nbytes:
            self.buf = None
            self.buf = alloc_workspace(nbytes, self.device)
        return self.buf

    def release(self):
        self.buf = None

    @property
    def nbytes(self):
        return 0 if self.buf is None else self.buf.numel()


class PWFSEngine:
    """Device-resident tables, workspace and the HIP calls for one PWFS geometry."""

    def __init__(self, n, N, cam, device=None, nchunk=None, pts_per_task=None, workspace=None):
        self.device = torch.device('cuda', torch.cuda.current_device()) if device is None else torch.device(device)
        self.n, self.N, self.cam = n, N, cam
        self.bin = N // cam if cam and N % cam == 0 else 0
        L = _lib.lib()
        self.fast_geometry = bool(L.oopao_pwfs_supported(n, N, self.bin))
        # a detector that does not divide N (non-integer pixels per subaperture): the separable transforms
        # give the unbinned frame, integrated onto the detector by area (oopao_bin_area)
        self.fractional = bool(cam) and N % cam != 0 and bool(L.oopao_pwfs_supported(n, N, 4))
        if not (self.fast_geometry or self.fractional) and not L.oopao_pwfs_generic_supported(n, N, self.bin):
            raise OopaoError('PWFS geometry n=%d N=%d binning=%d is not supported by the HIP path '
                             '(needs n <= N <= 2048 and a detector size dividing N, or N = 4 n on a compiled '
                             'length for a non-integer binning)' % (n, N, self.bin))
        self.mode = 'fast' if (self.fast_geometry or self.fractional) else 'generic'
        _lib.prepare_tables(N, self.device)
        self.nchunk_user = nchunk
        self.pts_per_task_user = pts_per_task
        # lanes per transform of the rows / columns pass (0 = the default plan; see plans())
        self.lanes_rows = 0
        self.lanes_cols = 0
        if self.mode == 'fast':
            rows, cols = self.plans()
            self._default_lanes = (rows[0], cols[0])
        else:
            self._default_lanes = (0, 0)
        # wavefronts the chip holds at the separable kernels' occupancy (three per SIMD, four SIMDs per CU)
        self._wave_slots = 12 * torch.cuda.get_device_properties(self.device).multi_processor_count \
            if self.device.type == 'cuda' else 3072
        self.workspace = Workspace(self.device) if workspace is None else workspace
        self.T = 0
        self.maskT = None
        self._fullT = None

    # ------------------------------------------------------------------ tables
    def set_modulation(self, path, tip_lin, tilt_lin, psfCentering=True):
        tx, ty = modulation_phasors(np.asarray(path, np.float64), tip_lin, tilt_lin, self.N, self.n,
                                    psfCentering)
        self.T = tx.shape[0]
        self._tx_host = tx
        self._ty_host = ty
        # general path: per-point tables in the reference's point order
        self.tx = torch.from_numpy(tx.astype(np.complex64)).to(self.device)
        self.ty = torch.from_numpy(ty.astype(np.complex64)).to(self.device)
        # fast path: distinct tip tables, points grouped by tip
        order, pidx, reps = tip_groups(path)
        self._order = order
        self.P = int(reps.size)
        self.tx_d = torch.from_numpy(tx[reps].astype(np.complex64)).to(self.device)
        pstart = np.searchsorted(pidx, np.arange(self.P + 1)).astype(np.int32)
        self.pstart = torch.from_numpy(pstart).to(self.device)
        self.ty_s = torch.from_numpy(ty[order].astype(np.complex64)).to(self.device)
        self.tx_s = torch.from_numpy(tx[order].astype(np.complex64)).to(self.device)   # per point, same order
        self._make_weighted()

    def set_weights(self, w):
        """Per-point weights: frames computed with weighted=True hold sum_i w_i I_i / T
        (Pyramid.py:750-755), by scaling tx_i with sqrt(w_i / T)."""
        self._weights = None if w is None else np.asarray(w, np.float64).reshape(-1)
        self._make_weighted()

    def _make_weighted(self):
        w = getattr(self, '_weights', None)
        self.txw = None
        self.ty_sw = None
        if w is None or self.T == 0:
            return
        if w.size != self.T:
            return      # reported when a weighted frame is asked for
        if np.any(w < 0):
            raise OopaoError('negative modulation weights are not supported')
        sc = np.sqrt(w / self.T)
        self.txw = torch.from_numpy((self._tx_host * sc[:, None]).astype(np.complex64)).to(self.device)
        tyw = self._ty_host[self._order] * sc[self._order][:, None]
        self.ty_sw = torch.from_numpy(tyw.astype(np.complex64)).to(self.device)

    def _weights_ready(self, weighted):
        w = getattr(self, '_weights', None)
        if not weighted or w is None:
            return False
        if self.txw is None:
            raise OopaoError('weight_vector has %d entries for %d modulation points' % (w.size, self.T))
        return True

    def _tx(self, weighted):
        return self.txw if self._weights_ready(weighted) else self.tx

    def _ty_sorted(self, weighted):
        return self.ty_sw if self._weights_ready(weighted) else self.ty_s

    def set_mask(self, m, mask=None, factor=None, general=False):
        """Pyramid mask from its phase m [N, N] (and the complex64 mask exp(i m) the reference
        stores, Pyramid.py:388). ``factor`` multiplies the mask in the focal plane (a spatial filter);
        ``general``: ``mask`` is not exp(i m) (an amplitude mask such as a Bi-O Edge knife edge) and is
        used as given. On a fast-path geometry a separable phase, or a given mask of rank one, selects the
        two-pass kernels; anything else the general path with the full mask."""
        m = np.asarray(m, np.float64)
        self._fullT = None
        f = None
        if (self.fast_geometry or self.fractional) and factor is None:
            if not general:
                s = separable_mask_factors(m)
                f = None if s is None else mask_vectors(s[0], s[1], self.N, self.n)
            elif mask is not None:
                r = rank1_mask_factors(mask)
                f = None if r is None else factor_vectors(r[0], r[1], self.N, self.n)
        if f is not None:
            my, mx = f
            self.my = torch.from_numpy(my.astype(np.complex64)).to(self.device)
            self.mx = torch.from_numpy(mx.astype(np.complex64)).to(self.device)
            self.mode = 'fast'
            self.maskT = None
            return
        if self.fractional:
            raise OopaoError('a non-integer detector sampling needs the separable pyramid mask (no rooftop, '
                             'rotation, per-quadrant shift, spatial filter or old mask)')
        full = np.complex64(np.exp(1j * m)) if mask is None else np.asarray(mask)
        full = full.astype(np.complex128)
        if factor is not None:
            full = full * np.asarray(factor)
        full = full / float(self.N) ** 2
        self.maskT = torch.from_numpy(np.ascontiguousarray(full.T).astype(np.complex64)).to(self.device)
        self.mode = 'generic'
        if self.bin == 0:
            raise OopaoError('detector size %d does not divide N=%d' % (self.cam, self.N))

    def plans(self):
        """FFT plans compiled for this N on the separable path, as lanes per transform:
        ([rows plans], [columns plans]), the default plan first."""
        r, c = (ctypes.c_int * 2)(), (ctypes.c_int * 2)()
        _lib.call('oopao_pwfs_plans', self.N, r, c)
        return [v for v in r if v], [v for v in c if v]

    def set_plans(self, lanes_rows=0, lanes_cols=0):
        """Select the rows / columns plans (lanes per transform, 0 = default); raises for a plan that is
        not compiled at this N."""
        rows, cols = self.plans()
        if (lanes_rows and lanes_rows not in rows) or (lanes_cols and lanes_cols not in cols):
            raise OopaoError('no compiled plan with %d / %d lanes at N=%d (rows %s, columns %s)'
                             % (lanes_rows, lanes_cols, self.N, rows, cols))
        self.lanes_rows, self.lanes_cols = int(lanes_rows), int(lanes_cols)

    # -------------------------------------------------------------- execution
    def _knobs(self, B, whole=False, fill=False):
        """Tuning defaults of the separable passes for B screens. Any value gives the same frames up to
        the float summation order over points; the chunk count does not depend on the column bands, so
        that a slopes-only frame keeps the whole frame's band columns bit for bit.

        Point chunks of the columns pass: about 9 points per chunk (config 2, one screen: 8 chunks,
        141-146 us slopes-only; 4-24 chunks within a few percent except 4); with ``fill`` (the binned
        columns pass), unless those chunks fill less than one round of the chip's wavefront slots: then
        the most chunks whose wavefronts, counted over every detector column, still fit in one round
        (config 1, one screen: 24 chunks, slopes-only 28 us against 64 us for 6 chunks, the whole frame
        33 us against 64 us, where 48 chunks, 1.25 rounds, take 42 us).

        Tip tables per rows-pass workgroup (the pass deals (table, row block) items out over the resident
        workgroups; the cap sets how many rounds): with 64-lane transforms each 128-byte line of the
        intermediate is completed by the four wavefronts of a workgroup, and workgroups that take several
        items let those drift apart and split the lines' writes, so the whole frame (twice the stores of
        a slopes-only one) takes one item per workgroup (config 2, 8 screens: 0.59 ms against 0.76 ms for
        2, one screen 68 us either way, 75-76 us for 4 or more) and slopes-only frames 4 (one screen: one
        round of 768 workgroups, 58 us against 60 us for 2; 8 screens 367 us against 375 us); with narrower
        transforms a wavefront writes whole lines itself, and 2 (whole) or 4 tables amortise the
        workgroup's start (config 1, 500 screens, whole: 1.35 ms against 1.45 ms for 1)."""
        tpt_r = self.lanes_rows or self._default_lanes[0]
        tpt_c = self.lanes_cols or self._default_lanes[1]
        if self.nchunk_user:
            nchunk = self.nchunk_user
        else:
            waves = lambda ppc: B * self.cam * (4 if tpt_c == 64 else 1) * -(-self.T // ppc)
            ppc = 9
            if fill and waves(ppc) < self._wave_slots:
                ppc = 1
                while ppc < 9 and waves(ppc) > self._wave_slots:
                    ppc += 1
            nchunk = -(-self.T // ppc)
        if self.pts_per_task_user:
            ppt = self.pts_per_task_user
        elif tpt_r == 64:
            ppt = 1 if whole else 4
        else:
            ppt = 2 if whole else 4
        return max(1, min(nchunk, self.T)), max(1, min(self.P, ppt))

    def _check(self, opd, amp):
        if self.T == 0:
            raise OopaoError('modulation tables not set')
        opd = opd.reshape(-1, self.n, self.n)
        if opd.dtype != torch.float32 or not opd.is_contiguous() or opd.device != self.device:
            raise OopaoError('opd must be a contiguous float32 tensor on %s' % self.device)
        if amp.dtype != torch.float32 or tuple(amp.shape) != (self.n, self.n) or not amp.is_contiguous():
            raise OopaoError('amp must be a contiguous float32 [n, n] tensor')
        return opd

    def _scratch(self, nbytes):
        return self.workspace.get(nbytes)

    def _generic_nchunk(self, B):
        """Point chunks of the general path: enough (screen, chunk, row) tasks to fill the GPU."""
        if self.nchunk_user:
            return min(self.nchunk_user, self.T)
        return max(1, min(self.T, -(-8192 // (B * self.N))))

    def bytes_per_screen(self):
        """Workspace bytes per phase screen of the current mode (for batch sizing)."""
        nb = ctypes.c_size_t(0)
        # the chunk count of a large batch (a small one may use more chunks, for parallelism)
        if self.mode == 'fast' and self.fractional:
            nchunk, _ = self._knobs(1, whole=True)
            _lib.call('oopao_pwfs_raw_bytes', self.n, self.N, self.T, 1, nchunk, ctypes.byref(nb))
            return nb.value + 4 * self.N * self.N
        if self.mode == 'fast':
            nchunk, _ = self._knobs(1, whole=True)
            _lib.call('oopao_pwfs_workspace_bytes', self.n, self.N, self.T, 1, nchunk, ctypes.byref(nb))
        else:
            _lib.call('oopao_pwfs_generic_bytes', self.n, self.N, self.T, 1, self._generic_nchunk(1),
                      ctypes.byref(nb))
        return nb.value

    def frames(self, opd, scale, amp, out=None, raw=None, weighted=False, bands=None):
        """Detector frames [B, cam, cam] float32 for OPD/phase maps opd [B, n, n] (phase = opd*scale).
        ``raw`` (a float32 [B, N, N] tensor) also receives the unbinned frames (the binned frames then
        come from them); ``weighted`` uses the per-point weights of set_weights. ``bands`` (separable path
        with 4 x 4 binning): detector-column intervals [(start, stop), ...] to compute, the other columns
        are returned as zero (the slopesMaps signal needs only the quadrant columns); ignored otherwise.
        With a non-integer binning (``fractional``) the unbinned frames are integrated by area."""
        opd = self._check(opd, amp)
        B = opd.shape[0]
        if out is None:
            out = torch.empty((B, self.cam, self.cam), dtype=torch.float32, device=self.device)
        if self.mode == 'fast' and (self.fractional or raw is not None):
            r = self._raw_fast(opd, scale, amp, weighted, raw)
            _lib.call('oopao_bin_area', _ptr(r), B, self.N, self.cam, _ptr(out), _stream_ptr(self.device))
            return out
        if self.mode == 'generic' or raw is not None:
            return self._frames_generic(opd, scale, amp, out, raw, self._tx(weighted))
        ty = self._ty_sorted(weighted)
        cb = column_bands(bands, self.cam)
        nchunk, ppt = self._knobs(B, whole=bands is None, fill=True)
        nb = ctypes.c_size_t(0)
        _lib.call('oopao_pwfs_workspace_bytes', self.n, self.N, self.T, B, nchunk, ctypes.byref(nb))
        ws = self._scratch(nb.value)
        _lib.call('oopao_pwfs_frames', _ptr(opd), ctypes.c_float(scale), _ptr(amp), _ptr(self.tx_d), self.P,
                  _ptr(self.pstart), _ptr(ty), _ptr(self.mx), _ptr(self.my), self.n, self.N, self.T, B, self.bin,
                  nchunk, ppt, self.lanes_rows, self.lanes_cols, cb, _ptr(ws),
                  ctypes.c_size_t(nb.value), _ptr(out), _stream_ptr(self.device))
        return out

    def _full_maskT(self):
        if self.maskT is None:
            if self._fullT is not None:
                return self._fullT
            # the fast path's separable mask as a full matrix (for the unbinned frame)
            mx = self.mx.cpu().numpy().astype(np.complex128)
            my = self.my.cpu().numpy().astype(np.complex128)
            c0 = self.N // 2 - self.n // 2
            k = np.arange(self.N)
            unshift = np.exp(2j * np.pi * k * c0 / self.N)
            full = np.outer(my * unshift, mx * unshift)
            self._fullT = torch.from_numpy(np.ascontiguousarray(full.T).astype(np.complex64)).to(self.device)
            return self._fullT
        return self.maskT

    def _frames_generic(self, opd, scale, amp, out, raw, tx):
        B = opd.shape[0]
        if raw is not None and (raw.dtype != torch.float32 or tuple(raw.shape) != (B, self.N, self.N)
                                or not raw.is_contiguous()):
            raise OopaoError('raw must be a contiguous float32 [B, N, N] tensor')
        nchunk = self._generic_nchunk(B)
        nb = ctypes.c_size_t(0)
        _lib.call('oopao_pwfs_generic_bytes', self.n, self.N, self.T, B, nchunk, ctypes.byref(nb))
        ws = self._scratch(nb.value)
        maskT = self._full_maskT()
        _lib.call('oopao_pwfs_generic_frames', _ptr(opd), ctypes.c_float(scale), _ptr(amp), _ptr(tx),
                  _ptr(self.ty), _ptr(maskT), self.n, self.N, self.T, B, self.bin, nchunk, _ptr(ws),
                  ctypes.c_size_t(nb.value), _ptr(out), ctypes.c_void_p(0 if raw is None else raw.data_ptr()),
                  _stream_ptr(self.device))
        return out

    def _raw_fast(self, opd, scale, amp, weighted, raw=None):
        """Unbinned frames [B, N, N] on the separable path (oopao_pwfs_raw)."""
        B = opd.shape[0]
        if raw is None:
            raw = torch.empty((B, self.N, self.N), dtype=torch.float32, device=self.device)
        elif raw.dtype != torch.float32 or tuple(raw.shape) != (B, self.N, self.N) or not raw.is_contiguous():
            raise OopaoError('raw must be a contiguous float32 [B, N, N] tensor')
        nchunk, ppt = self._knobs(B, whole=True)
        nb = ctypes.c_size_t(0)
        _lib.call('oopao_pwfs_raw_bytes', self.n, self.N, self.T, B, nchunk, ctypes.byref(nb))
        ws = self._scratch(nb.value)
        _lib.call('oopao_pwfs_raw', _ptr(opd), ctypes.c_float(scale), _ptr(amp), _ptr(self.tx_d), self.P,
                  _ptr(self.pstart), _ptr(self._ty_sorted(weighted)), _ptr(self.mx), _ptr(self.my), self.n, self.N,
                  self.T, B, nchunk, ppt, _ptr(ws), ctypes.c_size_t(nb.value), _ptr(raw), _stream_ptr(self.device))
        return raw

    def raw(self, opd, scale, amp, weighted=False):
        """Unbinned modulated frames [B, N, N] (the reference's raw_data, Pyramid.py:751-757)."""
        opd = self._check(opd, amp)
        if self.mode == 'fast':
            return self._raw_fast(opd, scale, amp, weighted)
        B = opd.shape[0]
        raw = torch.empty((B, self.N, self.N), dtype=torch.float32, device=self.device)
        out = torch.empty((B, self.cam, self.cam), dtype=torch.float32, device=self.device)
        self._frames_generic(opd, scale, amp, out, raw, self._tx(weighted))
        return raw

    def focal(self, opd, scale, amp, crop):
        """Modulation focal-plane intensity [B, crop, crop]: sum over points of |fft2(E_i)|^2 / N^2 on
        the window starting at N/2 - crop/2 (Pyramid.py:638-642, 1183-1190)."""
        opd = self._check(opd, amp)
        crop = int(crop)
        if not 0 < crop <= self.N:
            raise OopaoError('focal window %d outside 1..%d' % (crop, self.N))
        B = opd.shape[0]
        out = torch.empty((B, crop, crop), dtype=torch.float32, device=self.device)
        nb = ctypes.c_size_t(0)
        if self.fast_geometry or self.fractional:
            nchunk, ppt = self._knobs(B, whole=True)
            _lib.call('oopao_pwfs_focal_bytes', self.n, self.N, self.T, B, nchunk, crop, ctypes.byref(nb))
            ws = self._scratch(nb.value)
            _lib.call('oopao_pwfs_focal', _ptr(opd), ctypes.c_float(scale), _ptr(amp), _ptr(self.tx_d), self.P,
                      _ptr(self.pstart), _ptr(self.ty_s), self.n, self.N, self.T, B, nchunk, ppt, crop, _ptr(ws),
                      ctypes.c_size_t(nb.value), _ptr(out), _stream_ptr(self.device))
        else:
            nchunk = self._generic_nchunk(B)
            _lib.call('oopao_pwfs_generic_focal_bytes', self.n, self.N, self.T, B, nchunk, crop, ctypes.byref(nb))
            ws = self._scratch(nb.value)
            _lib.call('oopao_pwfs_generic_focal', _ptr(opd), ctypes.c_float(scale), _ptr(amp), _ptr(self.tx),
                      _ptr(self.ty), self.n, self.N, self.T, B, nchunk, crop, _ptr(ws), ctypes.c_size_t(nb.value),
                      _ptr(out), _stream_ptr(self.device))
        return out

    def adjoint(self, opd, scale, amp, gframe, weighted=False, bands=None):
        """dL/dopd [B, n, n] for the cotangent gframe = dL/dframes [B, cam, cam] of
        frames(opd, scale, amp, weighted=weighted): the separable path's adjoint (oopao_pwfs_adjoint) or the
        general path's (oopao_pwfs_generic_adjoint: any N, mask and integer binning); the modulation points
        are recomputed. ``bands`` (separable path, as in frames): detector-column intervals outside which
        gframe is zero (a loss of the slopesMaps signal); only those columns are transformed. Non-integer
        detector sampling (area-integrated pixels) has no adjoint."""
        if self.fractional:
            raise OopaoError('the PWFS adjoint does not cover non-integer detector sampling')
        opd = self._check(opd, amp)
        B = opd.shape[0]
        gframe = gframe.reshape(B, self.cam, self.cam)
        if gframe.dtype != torch.float32 or not gframe.is_contiguous() or gframe.device != self.device:
            gframe = gframe.to(self.device, torch.float32).contiguous()
        if self.mode == 'generic':
            return self._adjoint_generic(opd, scale, amp, gframe, weighted)
        # the backward kernels were tuned with 6 points per chunk
        nchunk = min(self.T, self.nchunk_user or max(1, -(-self.T // 6)))
        _, ppt = self._knobs(B, whole=bands is None)
        nb = ctypes.c_size_t(0)
        _lib.call('oopao_pwfs_adjoint_bytes', self.n, self.N, self.T, B, nchunk, ctypes.byref(nb))
        ws = self._scratch(nb.value)
        grad = torch.empty((B, self.n, self.n), dtype=torch.float32, device=self.device)
        _lib.call('oopao_pwfs_adjoint', _ptr(opd), ctypes.c_float(scale), _ptr(amp), _ptr(self.tx_d), self.P,
                  _ptr(self.pstart), _ptr(self._ty_sorted(weighted)), _ptr(self.tx_s), _ptr(self.mx), _ptr(self.my), _ptr(gframe),
                  self.n, self.N, self.T, B, nchunk, ppt, column_bands(bands, self.cam), _ptr(ws),
                  ctypes.c_size_t(nb.value), _ptr(grad), _stream_ptr(self.device))
        return grad

    def _adjoint_generic(self, opd, scale, amp, gframe, weighted):
        B = opd.shape[0]
        nchunk = self._generic_nchunk(B)
        nb = ctypes.c_size_t(0)
        _lib.call('oopao_pwfs_generic_adjoint_bytes', self.n, self.N, self.T, B, nchunk, ctypes.byref(nb))
        ws = self._scratch(nb.value)
        grad = torch.empty((B, self.n, self.n), dtype=torch.float32, device=self.device)
        _lib.call('oopao_pwfs_generic_adjoint', _ptr(opd), ctypes.c_float(scale), _ptr(amp), _ptr(self._tx(weighted)),
                  _ptr(self.ty), _ptr(self._full_maskT()), _ptr(gframe), self.n, self.N, self.T, B, self.bin, nchunk,
                  _ptr(ws), ctypes.c_size_t(nb.value), _ptr(grad), _stream_ptr(self.device))
        return grad

    # split passes (same result as frames(); used to time the passes separately)
    def buffers(self, B):
        nchunk, ppt = self._knobs(B, fill=True)
        ib, pb = ctypes.c_size_t(0), ctypes.c_size_t(0)
        _lib.call('oopao_pwfs_intermediate_bytes', self.n, self.N, self.T, B, nchunk, ctypes.byref(ib),
                  ctypes.byref(pb))
        inter = torch.empty(ib.value, dtype=torch.uint8, device=self.device)
        part = torch.empty(pb.value // 4, dtype=torch.float32, device=self.device)
        return dict(B=B, T=self.T, nchunk=nchunk, ppt=ppt, inter=inter, part=part)

    def _check_buffers(self, buf):
        if buf['T'] != self.T:
            raise OopaoError('PWFS buffers were sized for %d modulation points, the sensor has %d'
                             % (buf['T'], self.T))

    def rows(self, opd, scale, amp, buf, bands=None):
        self._check_buffers(buf)
        ppt = buf['ppt'] if bands is not None else self._knobs(buf['B'], whole=True)[1]
        _lib.call('oopao_pwfs_rows', _ptr(opd), ctypes.c_float(scale), _ptr(amp), _ptr(self.tx_d), self.P,
                  _ptr(self.pstart), _ptr(self.ty_s), _ptr(self.mx), self.n, self.N, self.T, buf['B'], ppt,
                  self.lanes_rows, column_bands(bands, self.cam), _ptr(buf['inter']),
                  ctypes.c_size_t(buf['inter'].numel()), _stream_ptr(self.device))

    def cols(self, buf, out, bands=None):
        self._check_buffers(buf)
        _lib.call('oopao_pwfs_cols', _ptr(buf['inter']), ctypes.c_size_t(buf['inter'].numel()), _ptr(self.my),
                  self.n, self.N, self.T, buf['B'], buf['nchunk'], self.lanes_cols, column_bands(bands, self.cam),
                  _ptr(buf['part']), ctypes.c_size_t(4 * buf['part'].numel()), _ptr(out), _stream_ptr(self.device))


class SlopesEngine:
    """Device tables for the slopesMaps reduction (Pyramid.py:835-854, 946-961)."""

    def __init__(self, device, cam, w, quad_origins):
        self.device = torch.device(device)
        self.cam = cam
        self.w = w
        self.quad = torch.tensor(np.asarray(quad_origins, np.int32).reshape(8), device=self.device)
        self.units = 1.0
        self._ws = None

    def set_valid(self, validI4Q):
        v = np.asarray(validI4Q, bool)
        self.validI4Q = v
        self.valid = torch.from_numpy(v.astype(np.uint8).reshape(-1)).to(self.device)
        vs = np.concatenate((v, v))
        self.vidx = torch.from_numpy(np.flatnonzero(vs.reshape(-1)).astype(np.int32)).to(self.device)
        self.nsig = int(self.vidx.numel())
        self.nvalid = int(v.sum())
        self.ref2d = torch.zeros((2 * self.w, self.w), dtype=torch.float32, device=self.device)

    def set_reference(self, ref2d):
        self.ref2d = ref2d.to(self.device, torch.float32).contiguous()

    def slopes(self, frames, sig2d=None, sig=None):
        frames = frames.reshape(-1, self.cam, self.cam)
        B = frames.shape[0]
        if sig2d is None:
            sig2d = torch.empty((B, 2 * self.w, self.w), dtype=torch.float32, device=self.device)
        if sig is None:
            sig = torch.empty((B, self.nsig), dtype=torch.float32, device=self.device)
        nb = ctypes.c_size_t(0)
        _lib.call('oopao_pwfs_slopes_workspace_bytes', B, self.w, ctypes.byref(nb))
        if self._ws is None or self._ws.numel() < nb.value:
            self._ws = torch.empty(nb.value, dtype=torch.uint8, device=self.device)
        _lib.call('oopao_pwfs_slopes', _ptr(frames), B, self.cam, _ptr(self.valid), self.w, _ptr(self.quad),
                  _ptr(self.ref2d), ctypes.c_float(self.units), _ptr(self.vidx), self.nsig, self.nvalid,
                  _ptr(self._ws), ctypes.c_size_t(nb.value), _ptr(sig2d), _ptr(sig), _stream_ptr(self.device))
        return sig2d, sig
