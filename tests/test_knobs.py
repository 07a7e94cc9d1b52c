"""CPU: the separable PWFS path's default tuning knobs (PWFSEngine._knobs).

The knobs change only the float summation order over points (the point chunks) or nothing at all (the
rows-pass tables per workgroup), so the properties that matter are checked here without a GPU: the chunk
count of a frame does not depend on its column bands (a slopes-only frame keeps the whole frame's band
columns bit for bit), the fill rule raises the chunk count only when the default chunks would leave the
chip's wavefront slots less than one round full, and user knobs override the defaults.
"""
import pytest

from oopao_amd.pwfs_engine import PWFSEngine


def _engine(n, T, P, lanes_rows, lanes_cols, slots=3072):
    e = PWFSEngine.__new__(PWFSEngine)     # the rules only: no tables, no device
    e.n, e.N, e.cam, e.T, e.P = n, 4 * n, n, T, P
    e.lanes_rows = e.lanes_cols = 0
    e._default_lanes = (lanes_rows, lanes_cols)
    e._wave_slots = slots
    e.nchunk_user = e.pts_per_task_user = None
    return e


@pytest.mark.parametrize('n,T,P,lr,lc', [(80, 48, 25, 16, 16), (320, 72, 37, 64, 64), (160, 48, 25, 32, 16),
                                         (256, 72, 37, 64, 64), (40, 24, 13, 16, 16)])
@pytest.mark.parametrize('B', [1, 2, 4, 8, 64, 500])
def test_chunks_do_not_depend_on_bands(n, T, P, lr, lc, B):
    e = _engine(n, T, P, lr, lc)
    assert e._knobs(B, whole=False, fill=True)[0] == e._knobs(B, whole=True, fill=True)[0]


def test_fill_rule():
    # config 2, one screen: 8 chunks of 9 points already fill more than one round (10240 wavefronts)
    assert _engine(320, 72, 37, 64, 64)._knobs(1, fill=True)[0] == 8
    # config 1, one screen: 6 chunks are 480 wavefronts; 24 chunks (1920) is the most that fits one round
    e = _engine(80, 48, 25, 16, 16)
    assert e._knobs(1, fill=False)[0] == 6
    assert e._knobs(1, fill=True)[0] == 24
    # a large batch keeps ~9 points per chunk
    assert e._knobs(500, fill=True)[0] == 6
    # the chosen chunks never exceed one round when the default did not fill it, and never exceed T
    for B in (1, 2, 3, 4, 5):
        nchunk = e._knobs(B, fill=True)[0]
        assert 1 <= nchunk <= e.T
        assert B * e.cam * nchunk <= max(e._wave_slots, B * e.cam * 6)


def test_rows_tables_per_workgroup():
    e64 = _engine(320, 72, 37, 64, 64)
    assert e64._knobs(1, whole=True)[1] == 1       # whole frames: one item per workgroup
    assert e64._knobs(1, whole=False)[1] == 4      # slopes-only, one screen
    assert e64._knobs(8, whole=False)[1] == 4      # slopes-only, several screens
    assert e64._knobs(8, whole=True)[1] == 1
    e16 = _engine(80, 48, 25, 16, 16)
    assert e16._knobs(1, whole=True)[1] == 2
    assert e16._knobs(64, whole=False)[1] == 4
    # capped by the number of tip tables
    assert _engine(16, 4, 1, 16, 16)._knobs(1, whole=False)[1] == 1


def test_user_knobs_override():
    e = _engine(80, 48, 25, 16, 16)
    e.nchunk_user, e.pts_per_task_user = 7, 3
    assert e._knobs(1, fill=True) == (7, 3)
    e.nchunk_user = 1000                            # clamped to T
    assert e._knobs(1, fill=True)[0] == 48
