"""CPU: the counter-based RNG of the atmosphere's device mode (oracle/rng.py) and the oracle's screen from
an arbitrary draw vector (oracle/atmosphere.screen_from_draws).

* Philox4x32-10 against the published known-answer vectors of the Random123 distribution (zero counter and
  key, all-ones counter and key, and the pi-digits vector).
* The normals: moments of 2^20 draws (mean, variance, fourth moment, lag-1 and pair correlations) within
  their sampling errors; draws, streams and keys give different numbers; a draw is a pure function of its
  coordinates (a prefix of a longer draw is the shorter draw).
* screen_from_draws with RandomState(seed)'s draws equals the seeded generators (both modes).
"""
import numpy as np
import pytest

from oracle import rng as orng


def test_philox_known_answers():
    def run(ctr, key):
        return [int(v) for v in orng.philox4x32(np.array(ctr, np.uint64), np.array(key, np.uint64))]
    assert run([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert run([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert run([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_normal_moments_and_independence():
    n = 1 << 20
    z = orng.normals(n, 1042, 0, 0, 7)
    se = 1 / np.sqrt(n)
    assert abs(z.mean()) < 5 * se
    assert abs(z.var() - 1) < 5 * np.sqrt(2) * se
    assert abs(np.mean(z ** 4) - 3) < 5 * np.sqrt(96) * se
    assert abs(np.mean(z[1:] * z[:-1])) < 5 * se           # lag 1 (inside and across Box-Muller pairs)
    assert abs(np.mean(z[0::2] * z[1::2])) < 5 * np.sqrt(2) * se
    # distinct coordinates, distinct numbers; a prefix is the shorter draw
    a = orng.normals(1001, 1042, 0, 0, 7)
    assert np.array_equal(a, z[:1001])
    for other in (orng.normals(1001, 1042, 0, 0, 8), orng.normals(1001, 1042, 0, 1, 7),
                  orng.normals(1001, 1043, 0, 0, 7), orng.normals(1001, 1042, 1, 0, 7)):
        assert abs(np.corrcoef(a, other)[0, 1]) < 0.15
    big = orng.normals(10, 5, 0, 0, (1 << 32) + 3)
    assert np.all(np.isfinite(big)) and not np.array_equal(big, orng.normals(10, 5, 0, 0, 3))


@pytest.mark.parametrize('M,mode', [(34, 2), (35, 2), (34, 1)])
def test_screen_from_draws_equals_seeded_generators(M, mode):
    from oracle import atmosphere as oat
    r0, L0, pix, seed = 0.15, 25.0, 0.1, 100017
    g = np.random.RandomState(seed).normal(size=2 * M * M)
    ref = (oat.ft_sh_phase_screen if mode == 2 else oat.ft_phase_screen)(r0, L0, M, pix, seed)
    got = oat.screen_from_draws(r0, L0, M, pix, g, mode)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
