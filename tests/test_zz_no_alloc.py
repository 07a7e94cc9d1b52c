"""GPU: the allocation-free slopes entry point (oopao_pwfs_slopes_static) against oopao_pwfs_slopes.

Both run the same kernels; the static variant takes its partial sums from a static device array
instead of a workspace argument. Two calls in a row (the second with fewer frames, so the array holds
the previous call's larger contents) must each equal the workspace entry point bit for bit, at the
config-1 (w = 20) and config-2 (w = 80) quadrant sizes; a call beyond the array's size is rejected
before any device work.
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _case(cam, w, seed):
    rs = np.random.RandomState(seed)
    e = w // 2
    c = cam // 2
    quad = np.array([c - e - w, c - e - w, c + e, c + e, c - e - w, c + e, c + e, c - e - w], np.int32)
    valid = (rs.rand(w, w) > 0.25)
    vs = np.concatenate((valid, valid))
    vidx = np.flatnonzero(vs.reshape(-1)).astype(np.int32)
    ref2d = (1e-3 * rs.randn(2 * w, w)).astype(np.float32)
    return quad, valid, vidx, ref2d


def _run(name, frames, cam, w, quad, valid, vidx, ref2d, units):
    from oopao_amd import _lib
    dev = frames.device
    B = frames.shape[0]
    q = torch.from_numpy(quad).to(dev)
    v = torch.from_numpy(valid.astype(np.uint8).reshape(-1)).to(dev)
    vi = torch.from_numpy(vidx).to(dev)
    r = torch.from_numpy(ref2d).to(dev)
    sig2d = torch.full((B, 2 * w, w), float('nan'), device=dev)
    sig = torch.full((B, vidx.size), float('nan'), device=dev)
    P = ctypes.c_void_p
    st = P(torch.cuda.current_stream(dev).cuda_stream)
    args = [P(frames.data_ptr()), B, cam, P(v.data_ptr()), w, P(q.data_ptr()), P(r.data_ptr()), ctypes.c_float(units),
            P(vi.data_ptr()), int(vidx.size), int(valid.sum())]
    if name == 'oopao_pwfs_slopes':
        nb = ctypes.c_size_t(0)
        _lib.call('oopao_pwfs_slopes_workspace_bytes', B, w, ctypes.byref(nb))
        ws = torch.empty(nb.value, dtype=torch.uint8, device=dev)
        args += [P(ws.data_ptr()), ctypes.c_size_t(nb.value)]
    _lib.call(name, *args, P(sig2d.data_ptr()), P(sig.data_ptr()), st)
    torch.cuda.synchronize()
    return sig2d, sig


@pytest.mark.parametrize('cam,w', [(80, 20), (320, 80)])
def test_static_slopes_equal_workspace_slopes(cam, w):
    dev = torch.device('cuda', 0)
    quad, valid, vidx, ref2d = _case(cam, w, seed=w)
    for call, B in enumerate((7, 2)):          # the second call is smaller than the first
        g = torch.Generator(device='cpu').manual_seed(100 * w + call)
        frames = (torch.rand((B, cam, cam), generator=g) * 1e4).to(dev)
        print('static slopes: cam=%d w=%d call %d B=%d' % (cam, w, call, B), flush=True)
        a2d, a = _run('oopao_pwfs_slopes_static', frames, cam, w, quad, valid, vidx, ref2d, 1.7)
        b2d, b = _run('oopao_pwfs_slopes', frames, cam, w, quad, valid, vidx, ref2d, 1.7)
        assert torch.equal(a2d, b2d) and torch.equal(a, b)
        assert not torch.isnan(a2d).any() and not torch.isnan(a).any()


def test_static_slopes_reject_oversize_call():
    from oopao_amd import _lib
    L = _lib.lib()
    w = 80
    nblk = (w * w + 255) // 256
    B = 65536 // nblk + 1
    dummy = ctypes.c_void_p(256)            # never dereferenced: the size check comes first
    rc = L.oopao_pwfs_slopes_static(dummy, B, 320, dummy, w, dummy, dummy, ctypes.c_float(1.0), dummy, 10, 10,
                                    dummy, dummy, None)
    assert rc == 4
    assert b'static scratch' in L.oopao_last_error()
