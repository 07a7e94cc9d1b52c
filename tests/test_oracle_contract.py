"""CPU: pins of the BF16 rounding oracle (oracle/contract.py)."""
import numpy as np
import torch

from oracle.contract import bf16_rne, split_bf16, product_terms
from tests.helpers import rel_l2


def test_bf16_rne_matches_torch():
    rs = np.random.RandomState(0)
    x = (rs.randn(20000) * np.exp(rs.uniform(-20, 20, 20000))).astype(np.float32)
    # exact ties: 8 significant bits plus a half unit in the last place, both parities
    m = rs.randint(128, 256, 2000).astype(np.float32)
    ties = np.ldexp(m + 0.5, rs.randint(-30, 30, 2000)).astype(np.float32) * np.where(rs.rand(2000) < .5, -1, 1)
    for v in (x, ties.astype(np.float32)):
        ref = torch.from_numpy(v).to(torch.bfloat16).float().numpy()
        assert np.array_equal(bf16_rne(v), ref)


def test_split_bounds():
    rs = np.random.RandomState(1)
    x = (rs.randn(50000) * np.exp(rs.uniform(-10, 10, 50000))).astype(np.float32)
    hi, lo = split_bf16(x)
    ax = np.abs(x.astype(np.float64))
    assert np.all(np.abs(x - hi.astype(np.float64)) <= 2.0 ** -8 * ax)
    assert np.all(np.abs(x - hi.astype(np.float64) - lo.astype(np.float64)) <= 2.0 ** -16 * ax)


def test_product_terms_bounds():
    rs = np.random.RandomState(2)
    A = rs.randn(60, 200).astype(np.float32)
    X = rs.randn(200, 7).astype(np.float32)
    exact = A.astype(np.float64) @ X.astype(np.float64)
    scale = np.abs(A.astype(np.float64)) @ np.abs(X.astype(np.float64))
    e1 = np.abs(product_terms(A, X, 1) - exact)
    e3 = np.abs(product_terms(A, X, 3) - exact)
    # bf16 keeps 8 significant bits (unit roundoff 2^-8): each hi*hi product is within
    # (2 * 2^-8 + 2^-16) of the exact one; the three-term form leaves the split's
    # representation errors (2 * 2^-16) and the dropped lo*lo (<= 2^-16 (1 + 2^-8)^2)
    assert np.all(e1 <= (2 * 2.0 ** -8 + 2.0 ** -16) * scale)
    assert np.all(e3 <= (3 * 2.0 ** -16 + 2.0 ** -20) * scale)
    assert rel_l2(product_terms(A, X, 3), exact) < rel_l2(product_terms(A, X, 1), exact) / 100
