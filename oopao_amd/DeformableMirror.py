"""DeformableMirror: Gaussian influence functions and OPD = modes @ coefs (DeformableMirror.py:25-735).

Geometry and influence functions follow the reference (306-340, 359-386,
572-600). The OPD is computed on the GPU by one of two exact routes:

* separable: for a Cartesian actuator grid without rotation/anamorphosis the
  Gaussian IF factorises, G_k(X, Y) = gx(X - x0_k) gy(Y - y0_k), so
  OPD = Gy C Gx^T with C the coefficient grid (k_sep_stage1/2 in dm.hip);
* dense: OPD = modes @ coefs with the [n^2, nAct] float32 modes resident in HBM
  (user-defined modes, rotated/anamorphic grids), one GEMV per column; with
  ``contraction = 'bf16'`` or ``'bf16x3'`` the modes are packed once to BF16
  (split BF16 for 'bf16x3') and the product runs on the matrix cores with FP32
  accumulation (contract.py); ``use_dense = True`` selects this route for any DM.

A wrong coefficient length raises OopaoError (the reference calls sys.exit(0),
DeformableMirror.py:683-685).
"""
import numpy as np
import torch

from . import _lib
from ._device import default_device, ptr, stream_ptr, to_host
from .MisRegistration import MisRegistration
from .tools import OopaoError


class DeformableMirror:
    def __init__(self, telescope, nSubap: float, mechCoupling: float = 0.35, coordinates: np.ndarray = None,
                 pitch: float = None, modes: np.ndarray = None, misReg=None, M4_param=None, nJobs: int = 30,
                 nThreads: int = 20, print_dm_properties: bool = True, floating_precision: int = 64,
                 altitude: float = None, flip=False, flip_lr=False, sign=1, actuator_selection=None,
                 user_defined_influence_functions_tag=None):
        if M4_param is not None:
            raise OopaoError('M4 influence functions are not supported by oopao_amd')
        if altitude is not None:
            raise OopaoError('altitude-conjugated DMs are not supported by oopao_amd')
        self.device = telescope.device
        self.print_dm_properties = print_dm_properties
        self.floating_precision = floating_precision
        self.flip_ = flip
        self.flip_lr = flip_lr
        self.sign = sign
        self.M4_param = None
        self.isM4 = False
        self.rad2arcsec = (180. / np.pi) * 3600
        self.actuator_selection = actuator_selection
        self.telescope = telescope
        self.altitude = None
        if mechCoupling <= 0:
            raise OopaoError('The value of mechanical coupling should be positive.')
        self.resolution = telescope.resolution
        self.mechCoupling = mechCoupling
        self.tag = 'deformableMirror'
        self.D = telescope.D
        self.pitch = self.D / nSubap if pitch is None else pitch
        self.misReg = MisRegistration() if misReg is None else misReg
        self._cartesian = coordinates is None
        if coordinates is None:
            # Cartesian geometry (DeformableMirror.py:306-340)
            self.nAct = nSubap + 1
            self.nActAlongDiameter = self.nAct - 1
            x = np.linspace(-self.D / 2, self.D / 2, self.nAct)
            X, Y = np.meshgrid(x, x)
            self.xIF0 = X.reshape(self.nAct ** 2)
            self.yIF0 = Y.reshape(self.nAct ** 2)
            r = np.sqrt(self.xIF0 ** 2 + self.yIF0 ** 2)
            if self.actuator_selection is None:
                r_in = telescope.centralObstruction * telescope.initial_D / 2 - 0.5 * self.pitch
                r_out = telescope.initial_D / 2 + 0.7533 * self.pitch
                self.actuator_selection = [r_in, r_out]
            if np.isscalar(self.actuator_selection):
                raise OopaoError('energy-based actuator selection is not supported by oopao_amd')
            self.validAct = (r >= self.actuator_selection[0]) * (r <= self.actuator_selection[1])
            self.nValidAct = int(np.sum(self.validAct))
        else:
            coordinates = np.asarray(coordinates)
            if coordinates.shape[1] != 2:
                raise OopaoError('Wrong size for the DM coordinates, the (x,y) coordinates should be input as '
                                 'a 2D array of dimension [nAct,2]')
            self.xIF0 = coordinates[:, 0]
            self.yIF0 = coordinates[:, 1]
            self.nAct = len(self.xIF0)
            self.nActAlongDiameter = self.D / self.pitch
            self.validAct = np.arange(self.nAct).astype(int)
            self.nValidAct = self.nAct
        xIF0 = self.xIF0[self.validAct]
        yIF0 = self.yIF0[self.validAct]
        self.nIF = len(xIF0)
        self.initial_coordinates = np.stack([xIF0, yIF0], axis=1)
        mr = self.misReg
        xIF3, yIF3 = self.anamorphosis_coordinates(xIF0, yIF0, mr.anamorphosisAngle * np.pi / 180,
                                                   mr.tangentialScaling, mr.radialScaling)
        xIF4, yIF4 = self.rotate_coordinates(xIF3, yIF3, mr.rotationAngle * np.pi / 180)
        self.xIF = xIF4 - mr.shiftX
        self.yIF = yIF4 - mr.shiftY
        self.u0x = self.resolution / 2 + self.xIF * self.resolution / self.D
        self.u0y = self.resolution / 2 + self.yIF * self.resolution / self.D
        self.coordinates = np.stack([self.xIF, self.yIF], axis=1)
        # Gaussian coefficients (DeformableMirror.py:573-588)
        cx = (1 + mr.radialScaling) * (self.resolution / self.nActAlongDiameter) / np.sqrt(2 * np.log(1. / mechCoupling))
        cy = (1 + mr.tangentialScaling) * (self.resolution / self.nActAlongDiameter) / np.sqrt(2 * np.log(1. / mechCoupling))
        theta = mr.anamorphosisAngle * np.pi / 180
        self._ga = np.cos(theta) ** 2 / (2 * cx ** 2) + np.sin(theta) ** 2 / (2 * cy ** 2)
        self._gb = -np.sin(2 * theta) / (4 * cx ** 2) + np.sin(2 * theta) / (4 * cy ** 2)
        self._gc = np.sin(theta) ** 2 / (2 * cx ** 2) + np.cos(theta) ** 2 / (2 * cy ** 2)
        self._modes_dev = None
        self._user_modes = modes is not None
        self._user_modes_f64 = None
        if modes is not None:
            m = np.asarray(modes) * self.sign
            self._user_modes_f64 = np.ascontiguousarray(m, dtype=np.float64)
            self._modes_dev = torch.from_numpy(np.ascontiguousarray(m, dtype=np.float32)).to(self.device)
            self.nValidAct = m.shape[1]
        self._sep = None if self._user_modes else self._separable_factors()
        self.use_dense = self._sep is None
        self.contraction = 'fp32'
        self._contract = None
        self._coefs = np.zeros(self.nValidAct)
        self.coefs = np.zeros(self.nValidAct)
        self.current_coefs = self.coefs.copy()
        if self.print_dm_properties:
            print(self)

    # ------------------------------------------------------------- geometry
    def rotate_coordinates(self, x, y, angle):
        return x * np.cos(angle) - y * np.sin(angle), y * np.cos(angle) + x * np.sin(angle)

    def anamorphosis_coordinates(self, x, y, angle, mRad, mNorm):
        mRad += 1
        mNorm += 1
        xOut = x * (mRad * np.cos(angle) ** 2 + mNorm * np.sin(angle) ** 2) + \
            y * (mNorm * np.sin(2 * angle) / 2 - mRad * np.sin(2 * angle) / 2)
        yOut = y * (mRad * np.sin(angle) ** 2 + mNorm * np.cos(angle) ** 2) + \
            x * (mNorm * np.sin(2 * angle) / 2 - mRad * np.sin(2 * angle) / 2)
        return xOut, yOut

    def _separable_factors(self):
        """gx [n, nX], gy [n, nY] and actuator grid indices when the IFs factorise, else None."""
        if self._gb != 0:
            return None
        ux = np.unique(np.round(self.u0x, 9))
        uy = np.unique(np.round(self.u0y, 9))
        if len(ux) * len(uy) > 4 * max(self.nValidAct, 1) + 64:
            return None  # not a grid
        ix = np.searchsorted(ux, np.round(self.u0x, 9))
        iy = np.searchsorted(uy, np.round(self.u0y, 9))
        n = self.resolution
        grid = np.linspace(0, 1, n) * n                       # DeformableMirror.py:580
        gx = np.exp(-self._ga * (grid[:, None] - ux[None, :]) ** 2) * self.sign
        gy = np.exp(-self._gc * (grid[:, None] - uy[None, :]) ** 2)
        if self.flip_lr:
            gx = gx[::-1]
        if self.flip_:
            gx = gx[::-1]
            gy = gy[::-1]
        # coefficient-grid scatter index: C[iy, ix] = coefs[k]
        nA_y, nA_x = len(uy), len(ux)
        if nA_x != nA_y:
            return None
        flat = iy * nA_x + ix
        grid = -np.ones(nA_x * nA_y, np.int32)
        grid[flat] = np.arange(flat.size, dtype=np.int32)
        return dict(nA=nA_x,
                    gxT=torch.from_numpy(np.ascontiguousarray(gx.T, np.float32)).to(self.device),
                    gy=torch.from_numpy(np.ascontiguousarray(gy, np.float32)).to(self.device),
                    grid=torch.from_numpy(grid).to(self.device))

    # ---------------------------------------------------------------- modes
    def _build_modes(self):
        n = self.resolution
        modes = torch.empty((n * n, self.nValidAct), dtype=torch.float32, device=self.device)
        x0 = torch.from_numpy(self.u0x.astype(np.float32)).to(self.device)
        y0 = torch.from_numpy(self.u0y.astype(np.float32)).to(self.device)
        _lib.call('oopao_dm_gaussian_modes', ptr(modes), n, self.nValidAct, ptr(x0), ptr(y0),
                  float(self._ga), float(self._gb), float(self._gc), float(self.sign), stream_ptr(self.device))
        if self.flip_lr or self.flip_:
            m = modes.reshape(n, n, -1)
            if self.flip_lr:
                m = torch.flip(m, dims=[1])
            if self.flip_:
                m = torch.flip(m, dims=[0, 1])
            modes = m.reshape(n * n, -1).contiguous()
        return modes

    def influence_functions_f64(self, chunk=512):
        """Influence functions [n^2, nValidAct] in float64 on the device: the Gaussian of
        DeformableMirror.py:572-600 evaluated in double precision (user modes as given)."""
        if self._user_modes:
            if self._user_modes_f64 is not None:
                return torch.from_numpy(self._user_modes_f64).to(self.device)
            return self._modes_dev.double()
        n = self.resolution
        grid = torch.from_numpy(np.linspace(0, 1, n) * n).to(self.device)
        Y, X = torch.meshgrid(grid, grid, indexing='ij')           # X varies along columns
        X = X.reshape(-1, 1)
        Y = Y.reshape(-1, 1)
        x0 = torch.from_numpy(np.asarray(self.u0x, np.float64)).to(self.device)
        y0 = torch.from_numpy(np.asarray(self.u0y, np.float64)).to(self.device)
        out = torch.empty((n * n, self.nValidAct), dtype=torch.float64, device=self.device)
        a, b, c = float(self._ga), float(self._gb), float(self._gc)
        for k0 in range(0, self.nValidAct, chunk):
            k1 = min(self.nValidAct, k0 + chunk)
            dx = X - x0[None, k0:k1]
            dy = Y - y0[None, k0:k1]
            out[:, k0:k1] = self.sign * torch.exp(-(a * dx * dx + 2 * b * dx * dy + c * dy * dy))
        if self.flip_lr or self.flip_:
            m = out.reshape(n, n, -1)
            if self.flip_lr:
                m = torch.flip(m, dims=[1])
            if self.flip_:
                m = torch.flip(m, dims=[0, 1])
            out = m.reshape(n * n, -1).contiguous()
        return out

    @property
    def modes(self):
        """Influence functions [n^2, nValidAct] (float64 numpy copy; DeformableMirror.py:391-394)."""
        if self._modes_dev is None:
            self._modes_dev = self._build_modes()
        return self._modes_dev.double().cpu().numpy()

    @modes.setter
    def modes(self, val):
        self._modes_dev = torch.from_numpy(np.ascontiguousarray(val, np.float32)).to(self.device)
        self._user_modes_f64 = np.ascontiguousarray(val, dtype=np.float64)
        self._user_modes = True
        self._sep = None
        self.use_dense = True
        self._contract = None

    # ---------------------------------------------------------------- coefs
    @property
    def coefs(self):
        return self._coefs

    @coefs.setter
    def coefs(self, val):
        """coefs setter (DeformableMirror.py:651-686): OPD = modes @ coefs on the GPU."""
        if np.isscalar(val):
            if val == 0:
                val = np.zeros(self.nValidAct)
            else:
                raise OopaoError('Error: wrong value for the coefficients')
        if isinstance(val, torch.Tensor):
            c = val.to(self.device, torch.float32)
            host = None
        else:
            host = np.asarray(val)
            c = torch.from_numpy(np.ascontiguousarray(host, dtype=np.float32)).to(self.device)
        if c.shape[0] != self.nValidAct:
            raise OopaoError('Error: wrong value for the coefficients (expected %d, got %d)'
                             % (self.nValidAct, c.shape[0]))
        self._coefs = host if host is not None else val
        self._OPD = self.opd_from_coefs(c)
        self.current_coefs = self._coefs.copy() if host is not None else val.clone()

    def opd_from_coefs(self, c):
        """Device OPD for coefficients c [nAct] or [nAct, B]: returns [n,n] or [B,n,n] float32."""
        n = self.resolution
        batched = c.dim() == 2
        cb = c.reshape(self.nValidAct, -1)
        B = cb.shape[1]
        if not self.use_dense:
            s = self._sep
            nA = s['nA']
            cc = cb.contiguous()
            opd = torch.empty((B, n, n), dtype=torch.float32, device=self.device)
            for j0 in range(0, B, 65535):
                j1 = min(B, j0 + 65535)
                cj = cc if (j0 == 0 and j1 == B) else cc[:, j0:j1].contiguous()
                tmp = torch.empty(((j1 - j0) * (nA * nA + n * nA),), dtype=torch.float32, device=self.device)
                _lib.call('oopao_dm_separable_opd', ptr(cj), ptr(s['grid']), ptr(s['gxT']), ptr(s['gy']), ptr(tmp),
                          ptr(opd[j0:j1]), n, nA, j1 - j0, stream_ptr(self.device))
        elif self.contraction != 'fp32':
            opd = self._bf16_contraction().gemm(cb, precision=self.contraction).reshape(B, n, n)
        else:
            if self._modes_dev is None:
                self._modes_dev = self._build_modes()
            opd = torch.empty((B, n * n), dtype=torch.float32, device=self.device)
            ct = cb.t().contiguous()
            for j in range(B):
                _lib.call('oopao_gemv', ptr(self._modes_dev), ptr(ct[j]), ptr(opd[j]), None, n * n,
                          self.nValidAct, 1.0, 0.0, stream_ptr(self.device))
            opd = opd.reshape(B, n, n)
        return opd if batched else opd[0]

    def _bf16_contraction(self):
        """Modes packed once to split BF16 MFMA fragments (contract.py), kept until the modes change."""
        from .contract import PRECISIONS, Contraction, PackedBF16
        if self.contraction not in PRECISIONS:
            raise OopaoError("contraction must be 'fp32', 'bf16' or 'bf16x3', got %r" % (self.contraction,))
        if self._contract is None:
            if self._modes_dev is None:
                self._modes_dev = self._build_modes()
            self._contract = Contraction(PackedBF16(self._modes_dev, split=True))
        return self._contract

    @property
    def OPD(self):
        return to_host(self._OPD)

    @OPD.setter
    def OPD(self, val):
        self._OPD = torch.from_numpy(np.ascontiguousarray(np.moveaxis(np.asarray(val), -1, 0)
                                                          if np.ndim(val) == 3 else val,
                                                          dtype=np.float32)).to(self.device)

    def set_coefs_value(self, actuator_index, actuator_coefs):
        if np.isscalar(actuator_index) and np.isscalar(actuator_coefs):
            actuator_index = [actuator_index]
            actuator_coefs = [actuator_coefs]
        if len(actuator_index) != len(actuator_coefs):
            raise OopaoError('The list of actuator_index and actuator_coefs must have the same length')
        c = np.array(self.coefs, dtype=np.float64)
        for i, v in zip(actuator_index, actuator_coefs):
            c[i] = v
        self.coefs = c

    # ---------------------------------------------------------------- relay
    def relay(self, src):
        """DeformableMirror.relay (DeformableMirror.py:427-457) on device tensors."""
        src_list = [src] if src.tag == 'source' else src.src
        for s in src_list:
            s.optical_path.append([self.tag, self])
            if s._OPD_no_pupil is None or s._OPD_no_pupil.dim() > 2:
                s._OPD_no_pupil = torch.zeros((self.resolution, self.resolution), dtype=torch.float32,
                                              device=self.device)
            if self._OPD.dim() == 2:
                s._set_OPD_no_pupil_dev(s._OPD_no_pupil + self._OPD)
            else:
                s._set_OPD_no_pupil_dev(self._OPD)

    def print_properties(self):
        print(self)

    def __repr__(self):
        return (f'\n------ Deformable mirror ------\nControlled Actuators     |{self.nValidAct:^9d}\n'
                f'Pitch [m]                |{self.pitch:^9.2f}\n'
                f'Mechnical coupling [%]   |{self.mechCoupling * 100:^9.0f}\n'
                f'OPD route                |{"dense" if self.use_dense else "separable":^9s}\n')
