"""BF16 matrix-core contractions for the dense influence functions and the reconstructor.

The reference computes the DM OPD as ``modes @ coefs`` (DeformableMirror.py:651-686,
float64) and applies the reconstructor as ``reconstructor @ wfs.signal``
(tutorials/AO_closed_loop_Pyramid_WFS.py:620-624). Here the resident matrix is
stored once in BF16 in the fragment layout of the gfx950 32x32x16 BF16 MFMA
(oopao_bf16_pack) and multiplied with FP32 accumulation:

* ``terms=1`` (``'bf16'``): hi * hi, about 3 significant decimal digits per product;
* ``terms=3`` (``'bf16x3'``, split BF16): Q = hi + lo with lo = bf16(Q - hi), and the
  product hi*hi + hi*lo + lo*hi, which keeps about 16 significant bits.

``gemm`` serves several right-hand sides (a batch of DM commands, a batch of slope
vectors) on the matrix cores; ``gemv`` serves one vector (HBM-bound: the BF16
matrix is half the bytes of float32).
"""
import ctypes

import torch

from . import _lib
from ._device import ptr, stream_ptr
from .tools import OopaoError

PRECISIONS = {'bf16': 1, 'bf16x3': 3}


def _packed_bytes(I, K):
    nb = ctypes.c_size_t(0)
    _lib.call('oopao_bf16_packed_bytes', I, K, ctypes.byref(nb))
    return nb.value


def _pack_into(src, si, sk, I, K, hi, lo, device):
    _lib.call('oopao_bf16_pack', ptr(src), si, sk, I, K, ptr(hi), None if lo is None else ptr(lo),
              stream_ptr(device))


class PackedBF16:
    """An I x K float32 matrix held as BF16 MFMA fragments: ``hi`` and, with ``split``, ``lo``.

    ``src`` is a float32 device tensor [I, K]; any strides are accepted (a transposed
    view packs the transpose).
    """

    def __init__(self, src, split=True):
        if not isinstance(src, torch.Tensor) or src.dim() != 2 or src.dtype != torch.float32:
            raise OopaoError('PackedBF16 needs a 2-D float32 device tensor')
        self.device = src.device
        self.I, self.K = int(src.shape[0]), int(src.shape[1])
        self.split = bool(split)
        nb = _packed_bytes(self.I, self.K)
        self.hi = torch.empty(nb // 2, dtype=torch.bfloat16, device=self.device)
        self.lo = torch.empty(nb // 2, dtype=torch.bfloat16, device=self.device) if split else None
        _pack_into(src, src.stride(0), src.stride(1), self.I, self.K, self.hi, self.lo, self.device)

    @property
    def nbytes(self):
        return self.hi.numel() * 2 * (2 if self.split else 1)

    def unpack(self, part='hi'):
        """The packed part as a float32 [I, K] tensor (inverse of the fragment layout)."""
        t = self.hi if part == 'hi' else self.lo
        if t is None:
            raise OopaoError('no lo part was packed')
        IT = (self.I + 31) // 32
        KTs = t.numel() // (IT * 512)     # stored k fragments (zero padding beyond K included)
        q = t.float().reshape(KTs, IT, 2, 32, 8).permute(1, 3, 0, 2, 4).reshape(IT * 32, KTs * 16)
        return q[:self.I, :self.K]


class Contraction:
    """Workspace-owning helper: out = alpha * A @ X for a resident PackedBF16 A."""

    def __init__(self, A: PackedBF16):
        self.A = A
        self._x = {}
        self._gemv_ws = None

    def _xbuf(self, B, split):
        key = (B, split)
        buf = self._x.get(key)
        if buf is None:
            nb = _packed_bytes(B, self.A.K)
            hi = torch.empty(nb // 2, dtype=torch.bfloat16, device=self.A.device)
            lo = torch.empty(nb // 2, dtype=torch.bfloat16, device=self.A.device) if split else None
            buf = self._x[key] = (hi, lo)
        return buf

    def gemm(self, X, precision='bf16x3', alpha=1.0, out=None):
        """out[b, r] = alpha * sum_k A[r, k] X[k, b] for X float32 [K, B] (or [K]); returns [B, R]."""
        terms = PRECISIONS.get(precision)
        if terms is None:
            raise OopaoError('precision must be one of %s, got %r' % (sorted(PRECISIONS), precision))
        if terms == 3 and not self.A.split:
            raise OopaoError('split BF16 needs the matrix packed with split=True')
        X = X.to(self.A.device, torch.float32)
        X2 = X.reshape(-1, 1) if X.dim() == 1 else X
        if X2.shape[0] != self.A.K:
            raise OopaoError('contraction length mismatch: matrix has K=%d, operand has %d'
                             % (self.A.K, X2.shape[0]))
        B = int(X2.shape[1])
        R = self.A.I
        xh, xl = self._xbuf(B, terms == 3)
        # Q[b][k] = X[k][b]
        _pack_into(X2, X2.stride(1), X2.stride(0), B, self.A.K, xh, xl, self.A.device)
        if out is None:
            out = torch.empty((B, R), dtype=torch.float32, device=self.A.device)
        elif out.shape != (B, R) or out.stride(1) != 1:
            raise OopaoError('out must be [B, R] with unit stride along R')
        _lib.call('oopao_gemm_bf16', ptr(self.A.hi), None if self.A.lo is None else ptr(self.A.lo), R, self.A.K,
                  ptr(xh), None if xl is None else ptr(xl), B, ptr(out), out.stride(0), float(alpha), terms,
                  stream_ptr(self.A.device))
        return out

    def gemv(self, x, y0=None, alpha=1.0, beta=0.0, out=None, precision='bf16'):
        """y = alpha * A @ x + beta * y0 with the BF16 matrix and an FP32 vector and accumulation:
        ``precision='bf16'`` multiplies A_hi (A to 2^-8), ``'bf16x3'`` (a split operand) adds A_lo @ x in
        a second pass (A = hi + lo to 2^-16; the same bytes as a float32 matrix)."""
        if precision not in PRECISIONS:
            raise OopaoError('precision must be one of %s, got %r' % (sorted(PRECISIONS), precision))
        if precision == 'bf16x3' and self.A.lo is None:
            raise OopaoError("precision 'bf16x3' needs a split operand (PackedBF16(..., split=True))")
        x = x.to(self.A.device, torch.float32).contiguous()
        if x.numel() != self.A.K:
            raise OopaoError('contraction length mismatch: matrix has K=%d, vector has %d' % (self.A.K, x.numel()))
        if self._gemv_ws is None:
            nb = ctypes.c_size_t(0)
            _lib.call('oopao_gemv_bf16_workspace_bytes', self.A.I, self.A.K, ctypes.byref(nb))
            self._gemv_ws = torch.empty(nb.value, dtype=torch.uint8, device=self.A.device)
        y = out if out is not None else torch.empty(self.A.I, dtype=torch.float32, device=self.A.device)
        first = y
        if precision == 'bf16x3':
            if getattr(self, '_gemv_tmp', None) is None:
                self._gemv_tmp = torch.empty(self.A.I, dtype=torch.float32, device=self.A.device)
            first = self._gemv_tmp
        _lib.call('oopao_gemv_bf16', ptr(self.A.hi), self.A.I, self.A.K, ptr(x), ptr(first),
                  None if y0 is None else ptr(y0), float(alpha), float(beta), ptr(self._gemv_ws),
                  self._gemv_ws.numel(), stream_ptr(self.A.device))
        if precision == 'bf16x3':
            # y = alpha * A_lo @ x + (alpha * A_hi @ x + beta * y0)
            _lib.call('oopao_gemv_bf16', ptr(self.A.lo), self.A.I, self.A.K, ptr(x), ptr(y), ptr(first),
                      float(alpha), 1.0, ptr(self._gemv_ws), self._gemv_ws.numel(), stream_ptr(self.A.device))
        return y
