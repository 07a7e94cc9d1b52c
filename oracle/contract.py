"""numpy definition of the BF16 (split) contractions checked by the tests (CPU oracle).

Test infrastructure only (see ``oracle/__init__.py``).

The reference product is float64 ``modes @ coefs`` (DeformableMirror.py:651-686) and
``reconstructor @ signal`` (tutorials/AO_closed_loop_Pyramid_WFS.py:620-624); the
float64 product itself is numpy's. What is restated here is the rounding of the
BF16 operands: round to nearest even of the float32 value to 8 significant bits
(``bf16_rne``), the split Q = hi + lo with lo = bf16(Q - hi) (``split_bf16``) and
the products the matrix cores form from them (``product_terms``).

Pins (tests/test_oracle_contract.py): ``bf16_rne`` against torch's float32 ->
bfloat16 conversion on random values and on exact ties; the split against its
defining bounds (|Q - hi| <= 2^-8 |Q|, |Q - hi - lo| <= 2^-16 |Q|); the
three-term product against the float64 product within the bound of the dropped
lo*lo term.
"""
import numpy as np


def bf16_rne(x):
    """float32 -> nearest BF16 (ties to even), returned as float32 values."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def split_bf16(x):
    """(hi, lo) with hi = bf16(x), lo = bf16(x - hi) (x - hi is exact in float32)."""
    x = np.asarray(x, dtype=np.float32)
    hi = bf16_rne(x)
    lo = bf16_rne(x - hi)
    return hi, lo


def product_terms(A, X, terms):
    """float64 value of the matrix-core product A @ X from the BF16 parts: terms = 1 -> hi hi;
    terms = 3 -> hi hi + hi lo + lo hi."""
    ah, al = (p.astype(np.float64) for p in split_bf16(A))
    xh, xl = (p.astype(np.float64) for p in split_bf16(X))
    if terms == 1:
        return ah @ xh
    return ah @ xh + ah @ xl + al @ xh
