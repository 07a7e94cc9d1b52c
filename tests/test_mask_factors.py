"""CPU: the mask factorisations that select the separable kernels (oopao_amd/pwfs_engine.py).

* separable_mask_factors: the default pyramid phase -c(|x| + |y|) (Pyramid.py:453-480) splits into
  a[k] + b[l]; a rooftop or rotated phase does not.
* rank1_mask_factors: the Bi-O Edge masks (BioEdge.py:331-390) without grey_length -- a knife-edge
  amplitude along one axis times a tip / tilt phase -- are p[k] q[l] to complex64 rounding; with
  grey_length (quadrant blocks) they are not; random rank-two masks are refused.
"""
import numpy as np

from oopao_amd.pwfs_engine import separable_mask_factors, rank1_mask_factors, factor_vectors, mask_vectors
from oopao_amd.BioEdge import bioedge_mask_set
from oracle import pwfs as opw


def test_pyramid_phase_split():
    g = opw.Geometry(80, 20, 20, 10)
    m, _ = opw.phase_mask(g)
    a, b = separable_mask_factors(m)
    assert np.max(np.abs(m - a[:, None] - b[None, :])) < 1e-9 * np.max(np.abs(m))
    for kw in (dict(rooftop=0.5), dict(theta_rotation=0.1)):
        m2, _ = opw.phase_mask(g, **kw)
        assert separable_mask_factors(m2) is None


def test_bioedge_masks_rank_one():
    N, n = 320, 80
    masks, _, _ = bioedge_mask_set(N, n, n, 20, N // n, grey_width=1, sx=(1, 0, -1, 0), sy=(0, 2, 0, -1))
    for M in masks:
        p, q = rank1_mask_factors(M)
        M = M.astype(np.complex128)
        assert np.max(np.abs(M - np.outer(p, q))) <= 2e-7 * np.max(np.abs(M))
        # the factors fold into the engine's vectors like a phase mask's: same shift, same 1/N
        my, mx = factor_vectors(p, q, N, n)
        c0 = N // 2 - n // 2
        unshift = np.exp(2j * np.pi * np.arange(N) * c0 / N) * N
        assert np.allclose(np.outer(my * unshift, mx * unshift), np.outer(p, q), atol=1e-12)
    masks_gl, _, _ = bioedge_mask_set(N, n, n, 20, N // n, grey_width=1, grey_length=2)
    assert all(rank1_mask_factors(M) is None for M in masks_gl)
    rs = np.random.RandomState(0)
    R2 = np.outer(rs.randn(N), rs.randn(N)) + np.outer(rs.randn(N), rs.randn(N))
    assert rank1_mask_factors(R2) is None


def test_mask_vectors_is_factor_vectors_of_exponentials():
    rs = np.random.RandomState(1)
    a, b = rs.randn(128), rs.randn(128)
    my1, mx1 = mask_vectors(a, b, 128, 32)
    my2, mx2 = factor_vectors(np.exp(1j * a), np.exp(1j * b), 128, 32)
    assert np.array_equal(my1, my2) and np.array_equal(mx1, mx2)
