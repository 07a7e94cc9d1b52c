"""Counter-based random numbers of the atmosphere's device RNG mode (CPU restatement).

Test infrastructure only (see ``oracle/__init__.py``).

* philox4x32: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
  SC 2011): a 4 x 32-bit counter and a 2 x 32-bit key, ten rounds of two 32 x 32 -> 64-bit multiplies
  (multipliers 0xD2511F53, 0xCD9E8D57) with the key bumped by the Weyl constants 0x9E3779B9, 0xBB67AE85.
  Pinned by the published known-answer vectors of the Random123 distribution.
* normals: the mapping of oopao_normal_philox (oopao_amd/csrc/rng.hip): pair p of draw d on stream s
  takes the Philox block of counter (p, d_lo, d_hi, s); two 53-bit uniforms u = (x0 >> 5, x1 >> 6) and
  (x2 >> 5, x3 >> 6), then Box-Muller z = sqrt(-2 ln(1 - u1)) (cos 2 pi u2, sin 2 pi u2).
"""
import numpy as np

M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox4x32(ctr, key, rounds=10):
    """Philox4x32 on arrays: ctr [..., 4] and key [..., 2] of uint32 values; returns [..., 4] uint64
    holding the 32-bit outputs."""
    c = [np.asarray(ctr[..., i], np.uint64) for i in range(4)]
    k0 = np.asarray(key[..., 0], np.uint64)
    k1 = np.asarray(key[..., 1], np.uint64)
    for r in range(rounds):
        p0 = c[0] * np.uint64(M0)
        p1 = c[2] * np.uint64(M1)
        hi0, lo0 = p0 >> np.uint64(32), p0 & np.uint64(MASK)
        hi1, lo1 = p1 >> np.uint64(32), p1 & np.uint64(MASK)
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
        if r < rounds - 1:
            k0 = (k0 + np.uint64(W0)) & np.uint64(MASK)
            k1 = (k1 + np.uint64(W1)) & np.uint64(MASK)
    return np.stack(c, axis=-1)


def normals(count, key0, key1, stream, draw):
    """``count`` float64 normals of draw ``draw`` on ``stream`` with key (key0, key1)."""
    npair = (count + 1) // 2
    p = np.arange(npair, dtype=np.uint64)
    ctr = np.stack([p, np.full(npair, draw & MASK, np.uint64), np.full(npair, (draw >> 32) & MASK, np.uint64),
                    np.full(npair, stream, np.uint64)], axis=-1)
    key = np.broadcast_to(np.array([key0, key1], np.uint64), (npair, 2))
    x = philox4x32(ctr, key)
    u1 = ((x[:, 0] >> np.uint64(5)).astype(np.float64) * 67108864.0 + (x[:, 1] >> np.uint64(6)).astype(np.float64)) \
        / 9007199254740992.0
    u2 = ((x[:, 2] >> np.uint64(5)).astype(np.float64) * 67108864.0 + (x[:, 3] >> np.uint64(6)).astype(np.float64)) \
        / 9007199254740992.0
    rad = np.sqrt(-2.0 * np.log(1.0 - u1))
    z = np.empty(2 * npair)
    z[0::2] = rad * np.cos(2 * np.pi * u2)
    z[1::2] = rad * np.sin(2 * np.pi * u2)
    return z[:count]


class PhiloxStream:
    """RandomState-like source of normals for the oracle layer: each normal(size) call is one draw of the
    stream, numbered from ``first``."""

    def __init__(self, key0, key1, stream, first=0):
        self.key0, self.key1, self.stream, self.draw = key0, key1, stream, first

    def normal(self, size):
        out = normals(int(np.prod(size)), self.key0, self.key1, self.stream, self.draw)
        self.draw += 1
        return out.reshape(size)
