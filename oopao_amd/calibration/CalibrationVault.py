"""CalibrationVault: SVD pseudo-inverse of an interaction matrix (calibration/CalibrationVault.py:14-123).

The SVD runs in float64 on the GPU (torch.linalg.svd -> rocSOLVER); the
attributes are float64 numpy arrays as in the reference, and ``M_dev`` keeps
the pseudo-inverse on the device.
"""
import numpy as np
import torch


class CalibrationVault:
    def __init__(self, D, nTrunc: int = 0, display: bool = False, print_details: bool = False,
                 invert: bool = True, device=None):
        if device is None:
            device = torch.device('cuda', torch.cuda.current_device()) if torch.cuda.is_available() else 'cpu'
        self.device = torch.device(device)
        Dt = D if isinstance(D, torch.Tensor) else torch.from_numpy(np.asarray(D, dtype=np.float64))
        Dt = Dt.to(self.device, torch.float64)
        if print_details:
            print('Computing the SVD...')
        if invert:
            U, s, Vh = torch.linalg.svd(Dt, full_matrices=False)
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
