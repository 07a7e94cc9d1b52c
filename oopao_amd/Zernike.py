"""Zernike: Noll-ordered Zernike modes on the telescope pupil (Zernike.py:10-147).

Host numpy, as in the reference (used at initialisation, e.g. for the specific
modes of the KL basis). Same normalisation: each mode except piston is made
zero-mean with unit standard deviation over the pupil.
"""
import math

import numpy as np


class Zernike:
    def __init__(self, telObject, J=1):
        self.resolution = telObject.resolution
        self.D = telObject.D
        self.centralObstruction = telObject.centralObstruction
        self.nModes = J

    @staticmethod
    def zernIndex(j):
        """(n, m) of Noll index j (Zernike.py:127-147)."""
        n = int((-1. + np.sqrt(8 * (j - 1) + 1)) / 2.)
        p = j - n * (n + 1) / 2.
        k = n % 2
        m = int((p + k) / 2.) * 2 - k
        if m != 0:
            m *= 1 if j % 2 == 0 else -1
        return [n, m]

    @staticmethod
    def zernikeRadialFunc(n, m, r):
        """Radial polynomial R_n^m(r) (Zernike.py:96-125)."""
        R = np.zeros(r.shape)
        for i in range(0, (n - m) // 2 + 1):
            coef = ((-1) ** i) * math.factorial(n - i) / (
                math.factorial(i) * math.factorial((n + m) // 2 - i) * math.factorial((n - m) // 2 - i))
            R += coef * r ** (n - 2 * i)
        return R

    def zernike_tel(self, tel, j, remove_piston):
        """Modes on the pupil pixels [npix, j] and full-resolution cube [n, n, j] (Zernike.py:17-66)."""
        pupil = np.asarray(tel.pupil)
        rows, cols = np.where(pupil > 0)
        off = (tel.resolution + tel.resolution % 2 - 1) / 2
        X = (rows - off) / tel.resolution * tel.D
        Y = (cols - off) / tel.resolution * tel.D
        R = np.sqrt(X ** 2 + Y ** 2)
        R = R / R.max()
        theta = np.arctan2(Y, X)
        npix = rows.size
        out = np.zeros([npix, j])
        full = np.zeros([tel.resolution ** 2, j])
        flat_idx = rows * tel.resolution + cols
        for i in range(remove_piston, j + remove_piston):
            n, m = self.zernIndex(i + 1)
            if m == 0:
                Z = np.sqrt(n + 1) * self.zernikeRadialFunc(n, 0, R)
            elif m > 0:
                Z = np.sqrt(2 * (n + 1)) * self.zernikeRadialFunc(n, m, R) * np.cos(m * theta)
            else:
                Z = np.sqrt(2 * (n + 1)) * self.zernikeRadialFunc(n, -m, R) * np.sin(-m * theta)
            if n != 0:
                Z = Z - Z.mean()
                Z = Z * (1 / np.std(Z))
            out[:, i - remove_piston] = Z
            full[flat_idx, i - remove_piston] = Z
        return out, full.reshape([tel.resolution, tel.resolution, j])

    def computeZernike(self, telObject2, remove_piston=1):
        self.modes, self.modesFullRes = self.zernike_tel(telObject2, self.nModes, remove_piston=remove_piston)

    def modeName(self, index):
        names = ['Tip', 'Tilt', 'Defocus', 'Astigmatism (X-shaped)', 'Astigmatism (+-shaped)',
                 'Coma vertical', 'Coma horizontal', 'Trefoil vertical', 'Trefoil horizontal',
                 'Sphere', 'Secondary astigmatism (X-shaped)', 'Secondary astigmatism (+-shaped)',
                 'Quadrofoil vertical', 'Quadrofoil horizontal',
                 'Secondary coma horizontal', 'Secondary coma vertical',
                 'Secondary trefoil horizontal', 'Secondary trefoil vertical',
                 'Pentafoil horizontal', 'Pentafoil vertical']
        if index < 0:
            return 'Incorrent index!'
        if index >= len(names):
            return ('Z', index + 2)
        return names[index]
