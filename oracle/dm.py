"""numpy restatement of OOPAO's Cartesian Gaussian deformable mirror (CPU oracle).

Test infrastructure only (see ``oracle/__init__.py``).

Pins (tests/test_oracle_dm.py): the valid-actuator count of the 21x21 / 81x81
grids against a direct count of the selection rule; the Gaussian against its
closed form at the actuator centre (1) and at one pitch (mechanical coupling);
the separable identity OPD = G_y C G_x^T, an independent route through the
same algebra.
"""
import numpy as np


class DMOracle:
    """Zero mis-registration Gaussian DM (DeformableMirror.py:306-340, 359-386, 572-600)."""

    def __init__(self, resolution, diameter, nSubap, mechCoupling=0.35, central_obstruction=0.0,
                 initial_diameter=None):
        n = resolution
        D = diameter
        D0 = diameter if initial_diameter is None else initial_diameter
        self.n = n
        self.nAct = nSubap + 1
        self.nActAlongDiameter = self.nAct - 1
        self.pitch = D / nSubap
        x = np.linspace(-D / 2, D / 2, self.nAct)
        X, Y = np.meshgrid(x, x)
        xIF0 = X.reshape(self.nAct ** 2)
        yIF0 = Y.reshape(self.nAct ** 2)
        r = np.sqrt(xIF0 ** 2 + yIF0 ** 2)
        r_in = central_obstruction * D0 / 2 - 0.5 * self.pitch
        r_out = D0 / 2 + 0.7533 * self.pitch
        self.validAct = (r >= r_in) * (r <= r_out)
        self.nValidAct = int(self.validAct.sum())
        self.xIF = xIF0[self.validAct]
        self.yIF = yIF0[self.validAct]
        self.u0x = n / 2 + self.xIF * n / D
        self.u0y = n / 2 + self.yIF * n / D
        self.c = (n / self.nActAlongDiameter) / np.sqrt(2 * np.log(1. / mechCoupling))
        self.grid = np.linspace(0, 1, n) * n           # DeformableMirror.py:580

    def modes(self, dtype=np.float64):
        """Influence functions [n^2, nValidAct] (DeformableMirror.py:572-600)."""
        n = self.n
        a = 1 / (2 * self.c ** 2)
        X, Y = np.meshgrid(self.grid, self.grid)
        out = np.empty((n * n, self.nValidAct), dtype=dtype)
        for k in range(self.nValidAct):
            G = np.exp(-(a * (X - self.u0x[k]) ** 2 + a * (Y - self.u0y[k]) ** 2))
            out[:, k] = G.reshape(n * n)
        return out

    def modes_rows(self, pixels):
        """Rows of the influence-function matrix for flat pupil pixels p = row * n + column, float64
        [len(pixels), nValidAct], from the Gaussian formula (without forming the whole matrix)."""
        p = np.asarray(pixels)
        a = 1 / (2 * self.c ** 2)
        X = self.grid[p % self.n][:, None]
        Y = self.grid[p // self.n][:, None]
        return np.exp(-(a * (X - self.u0x[None, :]) ** 2 + a * (Y - self.u0y[None, :]) ** 2))

    def opd(self, modes, coefs):
        """coefs setter (DeformableMirror.py:651-686): OPD = modes @ coefs, reshaped [n,n] or [n,n,B]."""
        coefs = np.asarray(coefs)
        out = np.matmul(modes, coefs)
        if coefs.ndim == 1:
            return out.reshape(self.n, self.n)
        return out.reshape(self.n, self.n, coefs.shape[1])

    def opd_separable(self, coefs):
        """The same OPD through the separable identity G_y C G_x^T (zero mis-registration Gaussian, pinned
        against modes @ coefs in tests/test_oracle_dm_atm.py); avoids the dense [n^2, nAct] modes."""
        ux = np.unique(np.round(self.u0x, 9))
        uy = np.unique(np.round(self.u0y, 9))
        a = 1 / (2 * self.c ** 2)
        gx = np.exp(-a * (self.grid[:, None] - ux[None, :]) ** 2)
        gy = np.exp(-a * (self.grid[:, None] - uy[None, :]) ** 2)
        C = np.zeros((len(uy), len(ux)))
        C[np.searchsorted(uy, np.round(self.u0y, 9)), np.searchsorted(ux, np.round(self.u0x, 9))] = coefs
        return gy @ C @ gx.T

