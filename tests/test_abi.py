"""CPU: the C-ABI library loads and exports every entry point declared in include/oopao_amd.h."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'oopao_amd.h')


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(oopao_[a-z0-9_]+)\s*\(', text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert 'oopao_pwfs_frames' in names and 'oopao_pwfs_slopes' in names
    assert len(names) >= 10


def test_library_exports_all_declared_symbols():
    from oopao_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip('library not built')
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers exactly the declared set
    assert set(_lib.SIGNATURES) == set(declared_functions())


def test_error_path_without_gpu_calls():
    from oopao_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip('library not built')
    L = _lib.lib()
    # argument validation happens before any device work
    assert L.oopao_pwfs_supported(320, 1280, 4) == 1
    assert L.oopao_pwfs_supported(320, 1000, 4) == 0
    nb = ctypes.c_size_t(0)
    assert L.oopao_pwfs_workspace_bytes(0, 1280, 72, 1, 8, ctypes.byref(nb)) != 0
    assert b'invalid' in L.oopao_last_error()
    assert L.oopao_pwfs_workspace_bytes(320, 1280, 72, 1, 73, ctypes.byref(nb)) != 0   # nchunk > T
    assert L.oopao_pwfs_workspace_bytes(320, 1280, 72, 1, 12, ctypes.byref(nb)) == 0
    assert nb.value >= 72 * 320 * 1280 * 8
    with pytest.raises(Exception):
        _lib.check(1)


def test_tip_groups_default_and_user_paths():
    # Pyramid.py:1129-1132: theta_i = 2 pi i / T, px = mod cos(theta): points i and T - i share their
    # tip, so the rows pass needs T/2 + 1 tables; a square user path has 3 distinct x values per side
    import numpy as np
    from oopao_amd.pwfs_engine import tip_groups
    from oracle.pwfs import modulation_path
    for T in (48, 72, 96):
        path = modulation_path(3, T)
        order, pidx, reps = tip_groups(path)
        assert reps.size == T // 2 + 1
        assert sorted(order.tolist()) == list(range(T))
        assert np.all(np.diff(pidx) >= 0) and pidx[-1] == reps.size - 1
        px = path[order, 0]
        assert np.max(np.abs(px - path[reps[pidx], 0])) < 1e-12
    sq = [[x, y] for x in (-1.0, 0.0, 1.0) for y in (-1.0, 0.0, 1.0) if (x, y) != (0.0, 0.0)]
    _, pidx, reps = tip_groups(np.asarray(sq))
    assert reps.size == 3


def test_fft_plan_query_without_gpu_calls():
    # oopao_pwfs_plans lists the compiled lane counts per pass, the default plan first (0 = none)
    from oopao_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip('library not built')
    L = _lib.lib()

    def plans(N):
        r, c = (ctypes.c_int * 2)(), (ctypes.c_int * 2)()
        rc = L.oopao_pwfs_plans(N, r, c)
        return rc, list(r), list(c)
    assert plans(1280) == (0, [64, 0], [64, 32])
    assert plans(320) == (0, [16, 32], [16, 32])
    assert plans(640) == (0, [32, 0], [16, 32])
    assert plans(1024) == (0, [64, 0], [64, 32])
    assert plans(2048) == (0, [64, 0], [64, 0])
    rc, _, _ = plans(1000)
    assert rc == 2 and b'N=1000' in L.oopao_last_error()
    assert L.oopao_pwfs_plans(1280, None, None) == 1


def test_generic_adjoint_sizes_and_argument_checks_without_gpu_calls():
    # workspace = pass-1 rows [B*T][n][N] + pass-2 field [B*T][N][N] (complex64) + chunk partial gradients
    # [B][nchunk][n][n] (float32), each rounded up to 256 bytes; invalid arguments fail before device work
    from oopao_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip('library not built')
    L = _lib.lib()
    al = lambda x: (x + 255) // 256 * 256
    nb = ctypes.c_size_t(0)
    n, N, T, B, nch = 120, 264, 20, 2, 5
    assert L.oopao_pwfs_generic_adjoint_bytes(n, N, T, B, nch, ctypes.byref(nb)) == 0
    assert nb.value == al(B * T * n * N * 8) + al(B * T * N * N * 8) + al(B * nch * n * n * 4)
    assert L.oopao_pwfs_generic_adjoint_bytes(n, N, T, B, T + 1, ctypes.byref(nb)) == 1
    assert L.oopao_pwfs_generic_adjoint_bytes(N + 1, N, T, B, nch, ctypes.byref(nb)) == 1
    d = ctypes.c_void_p(256)
    assert L.oopao_pwfs_generic_adjoint(None, 1.0, d, d, d, d, d, n, N, T, B, 6, nch, d, 1 << 40, d, None) == 1
    assert b'null' in L.oopao_last_error()
    # a binning that does not divide N, then a workspace smaller than required
    assert L.oopao_pwfs_generic_adjoint(d, 1.0, d, d, d, d, d, n, N, T, B, 7, nch, d, 1 << 40, d, None) == 1
    assert b'binning' in L.oopao_last_error()
    assert L.oopao_pwfs_generic_adjoint(d, 1.0, d, d, d, d, d, n, N, T, B, 6, nch, d, 100, d, None) == 4
