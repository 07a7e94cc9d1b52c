"""CalibrationVault: SVD pseudo-inverse of an interaction matrix (calibration/CalibrationVault.py:14-123).

The SVD runs in float64 on the GPU; the attributes are float64 numpy arrays
as in the reference, and ``M_dev`` keeps the pseudo-inverse on the device.

Two routes (``method``): 'svd' is torch.linalg.svd (rocSOLVER); 'gram' takes
the eigen-decomposition of the Gram matrix of the smaller side (D D^T or
D^T D, one DGEMM + a symmetric eigensolver, ~10x faster on the device at
632 x 1000 and 10080 x 1000) and recovers the other factor as
D^T U / s (or D V / s). The Gram route squares the condition number: the
pseudo-inverse then carries a relative error ~eps cond^2 and the singular
vectors of the smallest singular values ~eps cond^3. 'auto' (default) takes
it for matrices whose smaller side is >= 256 and whose condition number is
<= 1e4 (errors <= ~2e-8 on M, below the float32 noise of a measured
interaction matrix), and the SVD otherwise.
"""
import numpy as np
import torch

_GRAM_MIN_SIDE = 256
_GRAM_MAX_COND = 1e4


def thin_svd(D, method='auto'):
    """(U [m, r], s [r] descending, Vh [r, k]) of a float64 matrix D [m, k], r = min(m, k)."""
    if method not in ('auto', 'svd', 'gram'):
        raise ValueError("method must be 'auto', 'svd' or 'gram', got %r" % (method,))
    m, k = D.shape
    if method == 'svd' or (method == 'auto' and min(m, k) < _GRAM_MIN_SIDE):
        return torch.linalg.svd(D, full_matrices=False)
    wide = m <= k
    G = D @ D.t() if wide else D.t() @ D
    w, Q = torch.linalg.eigh(G)
    w, Q = w.flip(0), Q.flip(1)
    w_max, w_min = float(w[0]), float(w[-1])
    if w_min <= 0 or w_max / w_min > _GRAM_MAX_COND ** 2:
        if method == 'gram':
            raise ValueError('the Gram route needs a full-rank matrix with condition number <= %g' % _GRAM_MAX_COND)
        return torch.linalg.svd(D, full_matrices=False)
    s = torch.sqrt(w)
    if wide:
        return Q, s, (Q.t() @ D) / s[:, None]
    return (D @ Q) / s[None, :], s, Q.t()


class CalibrationVault:
    def __init__(self, D, nTrunc: int = 0, display: bool = False, print_details: bool = False,
                 invert: bool = True, device=None, method: str = 'auto'):
        if device is None:
            device = torch.device('cuda', torch.cuda.current_device()) if torch.cuda.is_available() else 'cpu'
        self.device = torch.device(device)
        Dt = D if isinstance(D, torch.Tensor) else torch.from_numpy(np.asarray(D, dtype=np.float64))
        Dt = Dt.to(self.device, torch.float64)
        if print_details:
            print('Computing the SVD...')
        if invert:
            U, s, Vh = thin_svd(Dt, method)
            self._U, self._s, self._V = U, s, Vh
            self.s = s.cpu().numpy()
            self.S = np.diag(self.s)
            self.eigenValues = self.s
            self.D = (U @ torch.diag(s) @ Vh).cpu().numpy()
            self.U = U.t().cpu().numpy()
            self.V = Vh.cpu().numpy()
            self.M_dev = Vh.t() @ torch.diag(1 / s) @ U.t()
            self.iS = np.diag(1 / self.eigenValues)
            self.M = self.M_dev.cpu().numpy()
            self._nTrunc = nTrunc
            self._truncate(nTrunc)
            if print_details:
                print('Done! The conditionning number is ' + str(self.cond))
        else:
            self.D = Dt.cpu().numpy()

    def _truncate(self, nTrunc):
        n = len(self.s) - nTrunc
        U, s, Vh = self._U, self._s, self._V
        self.iStrunc = np.diag(1 / self.eigenValues[:n])
        self.Vtrunc = self.V[:n, :]
        self.Utrunc = self.U[:n, :]
        self.VtruncT = self.Vtrunc.T
        self.UtruncT = self.Utrunc.T
        self.Mtrunc_dev = Vh[:n].t() @ torch.diag(1 / s[:n]) @ U[:, :n].t()
        self.Mtrunc = self.Mtrunc_dev.cpu().numpy()
        self.Dtrunc = (U[:, :n] @ torch.diag(s[:n]) @ Vh[:n]).cpu().numpy()
        self.cond = self.eigenValues[0] / self.eigenValues[(-nTrunc - 1)]

    @property
    def nTrunc(self):
        return self._nTrunc

    @nTrunc.setter
    def nTrunc(self, val):
        self._nTrunc = val
        self._truncate(val)
