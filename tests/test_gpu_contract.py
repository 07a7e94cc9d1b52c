"""GPU: BF16 matrix-core contractions (oopao_amd/csrc/contract.hip) against the float64 product.

Reference products: OPD = modes @ coefs (DeformableMirror.py:651-686) and
reconstructor @ signal (tutorials/AO_closed_loop_Pyramid_WFS.py:620-624).
Tolerances, stated per form:
* packing: bit-exact BF16 round-to-nearest-even hi and lo parts (oracle/contract.py);
* terms=1 (BF16): within 1e-5 relative of the float64 product of the BF16 operands
  (FP32 accumulation), and within the BF16 bound 2 * 2^-8 + 2^-16 of |A||X| of the
  exact product;
* terms=3 (split BF16): within 3 * 2^-16 + 2^-20 of |A||X| of the exact product plus
  FP32 accumulation (1e-6 of |A||X|);
* DM OPD through the split-BF16 route: rel-L2 <= 3e-5 against float64, and the PWFS
  slopes from it within rel-L2 1e-5 of the slopes from the float32 OPD.
"""
import numpy as np
import pytest
import torch

from oracle.contract import split_bf16, product_terms
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

DEV = 'cuda:0'


def _rand(shape, seed, scale=1.0):
    return (np.random.RandomState(seed).randn(*shape) * scale).astype(np.float32)


@pytest.mark.parametrize('I,K,transpose', [(37, 23, False), (300, 100, False), (129, 517, True)])
def test_pack_bit_exact(I, K, transpose):
    from oopao_amd.contract import PackedBF16
    print('pack', I, K, transpose, flush=True)
    Q = _rand((I, K), I + K)
    src = torch.from_numpy(np.ascontiguousarray(Q.T)).to(DEV).t() if transpose else torch.from_numpy(Q).to(DEV)
    P = PackedBF16(src, split=True)
    hi, lo = split_bf16(Q)
    assert np.array_equal(P.unpack('hi').cpu().numpy(), hi)
    assert np.array_equal(P.unpack('lo').cpu().numpy(), lo)
    # padding of the last fragments is zero
    assert int(torch.count_nonzero(P.hi.float())) == int(np.count_nonzero(hi))


@pytest.mark.parametrize('R,K,B,ldo_pad', [(300, 100, 37, 0), (1000, 517, 300, 5), (64, 16, 1, 0),
                                           (2050, 1030, 513, 3)])
def test_gemm_bf16(R, K, B, ldo_pad):
    from oopao_amd.contract import Contraction, PackedBF16
    print('gemm', R, K, B, ldo_pad, flush=True)
    A = _rand((R, K), 1)
    X = _rand((K, B), 2)
    C = Contraction(PackedBF16(torch.from_numpy(A).to(DEV), split=True))
    Xd = torch.from_numpy(X).to(DEV)
    exact = A.astype(np.float64) @ X.astype(np.float64)
    scale = np.abs(A.astype(np.float64)) @ np.abs(X.astype(np.float64))
    for terms, prec in ((1, 'bf16'), (3, 'bf16x3')):
        full = torch.full((B, R + ldo_pad), float('nan'), device=DEV)
        out = full[:, :R]
        C.gemm(Xd, precision=prec, alpha=0.5, out=out)
        got = out.double().cpu().numpy().T / 0.5
        if ldo_pad:
            assert torch.isnan(full[:, R:]).all()
        ref = product_terms(A, X, terms)
        err_ops = np.max(np.abs(got - ref) / scale)
        err = np.abs(got - exact)
        print(prec, 'max |err|/(|A||X|) vs operands %.2e' % err_ops, 'rel-L2 vs float64 %.3e' % rel_l2(got, exact))
        assert err_ops < 1e-5
        bound = (2 * 2.0 ** -8 + 2.0 ** -16) if terms == 1 else (3 * 2.0 ** -16 + 2.0 ** -20)
        assert np.all(err <= (bound + 1e-6) * scale)


def test_gemv_bf16():
    from oopao_amd.contract import Contraction, PackedBF16
    print('gemv', flush=True)
    R, K = 777, 1001
    A = _rand((R, K), 3)
    x = _rand((K,), 4)
    y0 = _rand((R,), 5)
    C = Contraction(PackedBF16(torch.from_numpy(A).to(DEV), split=False))
    y = C.gemv(torch.from_numpy(x).to(DEV), y0=torch.from_numpy(y0).to(DEV), alpha=-0.4, beta=1.0)
    hi = split_bf16(A)[0].astype(np.float64)
    ref = -0.4 * (hi @ x.astype(np.float64)) + y0
    scale = 0.4 * (np.abs(hi) @ np.abs(x.astype(np.float64))) + np.abs(y0)
    assert np.max(np.abs(y.double().cpu().numpy() - ref) / scale) < 1e-5
    exact = -0.4 * (A.astype(np.float64) @ x.astype(np.float64)) + y0
    print('gemv bf16 rel-L2 vs float64', rel_l2(y.double().cpu().numpy(), exact))
    assert np.all(np.abs(y.double().cpu().numpy() - exact) <= (2.0 ** -8 + 1e-5) * scale)


def test_dm_dense_bf16_routes_and_slopes():
    from oracle.dm import DMOracle
    from oopao_amd import Telescope, Source, DeformableMirror, Pyramid
    print('dm dense bf16', flush=True)
    tel = Telescope(80, 8, 1e-3, centralObstruction=0)
    ngs = Source('I3', 8, display_properties=False)
    ngs * tel
    dm = DeformableMirror(tel, 20, 0.35, print_dm_properties=False)
    wfs = Pyramid(20, tel, 3, 0.5, n_pix_separation=20, n_pix_edge=10, nTheta_user_defined=48)
    o = DMOracle(80, 8, 20, 0.35)
    C = np.random.RandomState(6).normal(0, 50e-9, (dm.nValidAct, 9))
    ref = o.opd(o.modes(), C)
    dm.use_dense = True
    errs = {}
    slopes = {}
    for prec in ('fp32', 'bf16', 'bf16x3'):
        dm.contraction = prec
        dm.coefs = C
        errs[prec] = rel_l2(dm.OPD, ref)
        tel.resetOPD()
        ngs * tel * dm * wfs
        slopes[prec] = np.array(wfs.signal)
    print('OPD rel-L2 vs float64', errs)
    assert errs['fp32'] < 1e-6
    assert errs['bf16x3'] < 3e-5
    assert errs['bf16'] < 5e-3
    s_err = {p: rel_l2(slopes[p], slopes['fp32']) for p in ('bf16', 'bf16x3')}
    print('slopes rel-L2 vs float32 route', s_err)
    assert s_err['bf16x3'] < 1e-5
