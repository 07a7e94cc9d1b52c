"""numpy restatement of OOPAO's frozen-flow atmosphere update (CPU oracle).

Test infrastructure only (see ``oracle/__init__.py``).

* warp_shift: the sub-pixel translation the reference delegates to
  scikit-image (tools/tools.py:202-209: warp(image, SimilarityTransform(
  translation=shift).inverse, order=3), mode 'constant', cval 0, clip=True),
  restated as Catmull-Rom cubic convolution (Keys, 1981, a = -1/2) on the 4x4
  neighbourhood with out-of-bounds samples equal to cval, followed by the
  clip of the result to [min, max] of the input. Pinned by
  tests/golden/warp_skimage.npz (outputs of scikit-image 0.18.3 itself).
* LayerOracle: updateLayer / add_row (Atmosphere.py:410-421, 447-516) and the
  pupil-footprint crop of fill_OPD_support (589-622) for an on-axis source at
  infinity, in float64, consuming the layer's RandomState exactly like the
  reference (one normal(size=nOuter) draw per extruded ring); with a boiling
  time, apply_boiling (387-408) first: the AR(1) mix of mapShift with a fresh
  ft_sh_phase_screen / ft_phase_screen of seed boiling_seed + 1.
* ft_phase_screen, ft_sh_phase_screen: phaseStats.py:153-271 (numpy FFTs, the
  reference's RandomState draw order: the subharmonics restart the seed).
* warp_bilinear: scikit-image's order-1 warp (mode 'constant', cval 0) for an
  affine map y = a r + by, x = a c + bx: the four floor/ceil neighbours of the
  mapped coordinate, any neighbour outside the image counted as cval. Pinned by
  tests/golden/bilinear_skimage.npz (scikit-image 0.18.3 itself, on the
  translation and scale-about-centre transforms the reference composes in
  tools/interpolateGeometricalTransformation.py:152-265).
* source_layer: one layer's contribution to an off-axis / laser-guide-star
  source (set_pupil_footprint, Atmosphere.py:423-445; fill_OPD_support,
  589-622): footprint position, bilinear sub-pixel translation of the whole
  layer, crop, cone magnification.
* asm, scintillation_field: the angular-spectrum propagation (Atmosphere.ASM,
  938-972, equal pitches) and the top-down layer loop of
  fill_scintillation_support (644-662). Pinned by the analytic Fresnel
  propagation of a Gaussian beam, Parseval and the zero-distance identity.
"""
import numpy as np


def _cubic(x, f0, f1, f2, f3):
    return f1 + 0.5 * x * (f2 - f0 + x * (2.0 * f0 - 5.0 * f1 + 4.0 * f2 - f3 + x * (3.0 * (f1 - f2) + f3 - f0)))


def warp_shift(img, tx, ty, cval=0.0, clip=True):
    """out[r, c] = cubic interpolation of img at (r - ty, c - tx)."""
    img = np.asarray(img, dtype=np.float64)
    R, C = img.shape
    rr, cc = np.meshgrid(np.arange(R, dtype=np.float64) - ty, np.arange(C, dtype=np.float64) - tx, indexing='ij')
    r0 = np.floor(rr).astype(np.int64)
    c0 = np.floor(cc).astype(np.int64)
    xr = rr - r0
    xc = cc - c0

    def px(r, c):
        ok = (r >= 0) & (r < R) & (c >= 0) & (c < C)
        v = np.full(r.shape, cval, dtype=np.float64)
        v[ok] = img[r[ok], c[ok]]
        return v

    rows = []
    for pr in range(-1, 3):
        f = [px(r0 + pr, c0 + pc) for pc in range(-1, 3)]
        rows.append(_cubic(xc, *f))
    out = _cubic(xr, *rows)
    if clip:
        lo, hi = img.min(), img.max()
        preserve = not (lo <= cval <= hi)
        if preserve:
            mask = out == cval
        out = np.clip(out, lo, hi)
        if preserve:
            out[mask] = cval
    return out


def ft_phase_screen(r0, L0, N, delta, seed, l0=1e-10):
    """phaseStats.ft_phase_screen (phaseStats.py:153-198), radians at 500 nm."""
    rs = np.random.RandomState(seed)
    del_f = 1. / (N * delta)
    fx = np.arange(-N / 2., N / 2.) * del_f
    fx, fy = np.meshgrid(fx, fx)
    f = np.sqrt(fx ** 2 + fy ** 2)
    fm = 5.92 / l0 / (2 * np.pi)
    psd = 0.023 * r0 ** (-5. / 3.) * np.exp(-1 * ((f / fm) ** 2)) / (((f ** 2) + (1. / L0) ** 2) ** (11. / 6))
    psd[int(N / 2), int(N / 2)] = 0
    cn = (rs.normal(size=(N, N)) + 1j * rs.normal(size=(N, N))) * np.sqrt(psd) * del_f
    return np.fft.fftshift(np.fft.fft2(np.fft.fftshift(cn))).real


def ft_sh_phase_screen(r0, L0, resolution, pixel_size, seed, l0=1e-10):
    """phaseStats.ft_sh_phase_screen (phaseStats.py:201-271): FFT screen + 3 levels of the 2 x 2
    subharmonics of a 3 x 3 grid, the subharmonic draws from a second RandomState(seed)."""
    rs = np.random.RandomState(seed)
    D = resolution * pixel_size
    hi = ft_phase_screen(r0, L0, resolution, pixel_size, seed, l0)
    coords = np.arange(-resolution / 2, resolution / 2) * pixel_size
    x, y = np.meshgrid(coords, coords)
    lo = np.zeros((resolution, resolution), complex)
    for p in range(1, 4):
        del_f = 1 / (3 ** p * D)
        fx = np.arange(-1, 2) * del_f
        fx, fy = np.meshgrid(fx, fx)
        f = np.sqrt(fx ** 2 + fy ** 2)
        fm = 5.92 / l0 / (2 * np.pi)
        psd = 0.023 * r0 ** (-5. / 3) * np.exp(-1 * (f / fm) ** 2) / ((f ** 2 + (1. / L0) ** 2) ** (11. / 6))
        psd[1, 1] = 0
        cn = (rs.normal(size=(3, 3)) + 1j * rs.normal(size=(3, 3))) * np.sqrt(psd) * del_f
        for i in range(0, 2):
            for j in range(0, 2):
                lo += cn[i, j] * np.exp(1j * 2 * np.pi * (fx[i, j] * x + fy[i, j] * y))
    return (lo.real - lo.real.mean()) + hi


def screen_from_draws(r0, L0, M, pixel_size, g, mode=2, l0=1e-10):
    """The phase screen ft_sh_phase_screen (mode 2) or ft_phase_screen builds from its normal draws g
    (2 M^2 values: real parts, then imaginary parts of the FFT-grid coefficients; the subharmonics take
    the first 54), phaseStats.py:153-271; with g = RandomState(seed).normal(size=2 M^2) it is the seeded
    screen."""
    g = np.asarray(g, np.float64)
    del_f = 1. / (M * pixel_size)
    fx = np.arange(-M / 2., M / 2.) * del_f
    fx, fy = np.meshgrid(fx, fx)
    f = np.sqrt(fx ** 2 + fy ** 2)
    fm = 5.92 / l0 / (2 * np.pi)
    psd = 0.023 * r0 ** (-5. / 3.) * np.exp(-1 * ((f / fm) ** 2)) / (((f ** 2) + (1. / L0) ** 2) ** (11. / 6))
    psd[int(M / 2), int(M / 2)] = 0
    cn = (g[:M * M] + 1j * g[M * M:2 * M * M]).reshape(M, M) * np.sqrt(psd) * del_f
    hi = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(cn))).real
    if mode != 2:
        return hi
    D = M * pixel_size
    coords = np.arange(-M / 2, M / 2) * pixel_size
    x, y = np.meshgrid(coords, coords)
    lo = np.zeros((M, M), complex)
    sub = g[:54].reshape(3, 2, 3, 3)
    for p in range(1, 4):
        d = 1 / (3 ** p * D)
        fxs, fys = np.meshgrid(np.arange(-1, 2) * d, np.arange(-1, 2) * d)
        fs = np.sqrt(fxs ** 2 + fys ** 2)
        ps = 0.023 * r0 ** (-5. / 3) * np.exp(-1 * (fs / fm) ** 2) / ((fs ** 2 + (1. / L0) ** 2) ** (11. / 6))
        ps[1, 1] = 0
        c = (sub[p - 1, 0] + 1j * sub[p - 1, 1]) * np.sqrt(ps) * d
        for i in range(2):
            for j in range(2):
                lo += c[i, j] * np.exp(1j * 2 * np.pi * (fxs[i, j] * x + fys[i, j] * y))
    return (lo.real - lo.real.mean()) + hi


class LayerOracle:
    """One frozen-flow layer: mapShift [(res+2)^2] float64, A, B, RandomState (Atmosphere.py:281-374, 410-516)."""

    def __init__(self, mapShift, A, B, inner_mask, outer_mask, random_state, ratio):
        self.mapShift = np.array(mapShift, dtype=np.float64)
        self.A = np.asarray(A, np.float64)
        self.B = np.asarray(B, np.float64)
        self.inner = np.asarray(inner_mask) != 0        # [(res+2)^2] ring masks
        self.outer = np.asarray(outer_mask) != 0
        self.rs = random_state
        self.ratio = np.asarray(ratio, np.float64)
        self.buff = np.zeros(2)
        self.res = self.mapShift.shape[0] - 2
        self.OPD = self.mapShift[1:-1, 1:-1].copy()
        self.boiling = None

    def set_boiling(self, t_boiling, sampling_time, r0, L0, pixel_size, boiling_seed, mode=2, draws=None):
        """Boiling state of the layer (Atmosphere.py:329-333, 387-408). ``draws(seed, count)`` gives a
        screen's normal draws (default: RandomState(seed), the reference's)."""
        self.boiling = dict(t=t_boiling, dt=sampling_time, r0=r0, L0=L0, pixel=pixel_size, mode=mode, draws=draws)
        self.boiling_seed = boiling_seed

    def apply_boiling(self):
        """Atmosphere.apply_boiling (Atmosphere.py:387-408): mapShift = alpha mapShift + beta phi."""
        b = self.boiling
        if b is None or b['t'] is None or not np.isfinite(b['t']) or b['t'] <= 0:
            return False
        alpha = float(np.exp(-b['dt'] / b['t']))
        if alpha >= 1.0 - 1e-6:
            return False
        beta = float(np.sqrt(max(0.0, 1.0 - alpha ** 2)))
        self.boiling_seed += 1
        M = self.mapShift.shape[0]
        if b['draws'] is None:
            gen = ft_sh_phase_screen if b['mode'] == 2 else ft_phase_screen
            phi = gen(b['r0'], b['L0'], M, b['pixel'], seed=self.boiling_seed)
        else:
            g = b['draws'](self.boiling_seed, 2 * M * M)
            phi = screen_from_draws(b['r0'], b['L0'], M, b['pixel'], g, b['mode'])
        self.mapShift = alpha * self.mapShift + beta * phi
        return True

    def add_row(self, step):
        """Atmosphere.add_row (Atmosphere.py:410-421)."""
        tmp = warp_shift(self.mapShift, step[0], step[1])
        one = tmp[1:-1, 1:-1]
        Z = one[self.inner[1:-1, 1:-1]]
        X = self.A @ Z + self.B @ self.rs.normal(size=self.B.shape[1])
        self.mapShift[self.outer] = X
        self.mapShift[~self.outer] = one.reshape(self.res * self.res)
        return one

    def update(self):
        """Atmosphere.updateLayer (Atmosphere.py:447-516): boiling first, then the frozen-flow shift (a
        layer without wind delivers the boiled support's interior)."""
        boiled = self.apply_boiling()
        ratio = self.ratio
        if ratio[0] == 0 and ratio[1] == 0:
            if boiled:
                self.OPD = self.mapShift[1:-1, 1:-1].copy()
            return self.OPD
        tmp = np.abs(ratio)
        tmp[np.isinf(tmp)] = 0
        nScreens = tmp.astype('int')
        step = np.zeros(2)
        sub = np.zeros(2)
        for _ in range(nScreens.min()):
            step[0] = 1
            step[1] = 1
            step = step * np.sign(ratio)
            self.OPD = self.add_row(step)
        for _ in range(nScreens.max() - nScreens.min()):
            step[0] = 1
            step[1] = 1
            step = step * np.sign(ratio)
            step[np.where(nScreens == nScreens.min())] = 0
            self.OPD = self.add_row(step)
        sub[0] = (np.abs(ratio[0]) % 1) * np.sign(ratio[0])
        sub[1] = (np.abs(ratio[1]) % 1) * np.sign(ratio[1])
        self.buff += sub
        if np.abs(self.buff[0]) >= 1 or np.abs(self.buff[1]) >= 1:
            step[0] = 1 * np.sign(self.buff[0])
            step[1] = 1 * np.sign(self.buff[1])
            step[np.where(np.abs(self.buff) < 1)] = 0
            self.OPD = self.add_row(step)
        self.buff[0] = (np.abs(self.buff[0]) % 1) * np.sign(self.buff[0])
        self.buff[1] = (np.abs(self.buff[1]) % 1) * np.sign(self.buff[1])
        self.OPD = warp_shift(self.mapShift, self.buff[0], self.buff[1])[1:-1, 1:-1]
        return self.OPD


def footprint_crop(layer_opd, n):
    """On-axis pupil footprint of a layer map (Atmosphere.py:309-315, 614-616)."""
    c = layer_opd.shape[0] // 2
    return layer_opd[c - n // 2:c + n // 2, c - n // 2:c + n // 2]


def atmosphere_opd(layers, fractionalR0, n, wavelength=500e-9):
    """Sum of layer OPDs in metres on the pupil (fill_OPD_support, Atmosphere.py:589-622)."""
    out = np.zeros((n, n))
    for lay, f in zip(layers, fractionalR0):
        out += footprint_crop(lay.OPD, n) * wavelength / 2 / np.pi * np.sqrt(f)
    return out


def warp_bilinear(img, a, by, bx, out_shape, cval=0.0):
    """out[r, c] = bilinear interpolation of img at (a r + by, a c + bx), neighbours outside = cval."""
    img = np.asarray(img, dtype=np.float64)
    H, W = img.shape
    rr, cc = np.meshgrid(np.arange(out_shape[0]) * a + by, np.arange(out_shape[1]) * a + bx, indexing='ij')
    r0, c0 = np.floor(rr), np.floor(cc)
    r1, c1 = np.ceil(rr), np.ceil(cc)
    dr, dc = rr - r0, cc - c0

    def px(r, c):
        ok = (r >= 0) & (r <= H - 1) & (c >= 0) & (c <= W - 1)
        v = np.full(r.shape, cval)
        v[ok] = img[r[ok].astype(np.int64), c[ok].astype(np.int64)]
        return v
    top = (1 - dc) * px(r0, c0) + dc * px(r0, c1)
    bot = (1 - dc) * px(r1, c0) + dc * px(r1, c1)
    return (1 - dr) * top + dr * bot


def source_layer(layer_full, n, layer_D, layer_altitude, zenith_arcsec, azimuth_deg, src_altitude,
                 chromatic_shift=0.0):
    """The layer's phase (radians at 500 nm) on the footprint of a source (Atmosphere.py:423-445, 589-622)."""
    res = layer_full.shape[0]
    rad2arcsec = 180 / np.pi * 3600
    rho = layer_altitude * np.tan((zenith_arcsec + chromatic_shift) / rad2arcsec) * res / layer_D
    x_z, y_z = rho * np.cos(np.deg2rad(azimuth_deg)), rho * np.sin(np.deg2rad(azimuth_deg))
    esx, esy = int(x_z) - x_z, int(y_z) - y_z
    im = np.asarray(layer_full, dtype=np.float64)
    if esx != 0 or esy != 0:
        # interpolate_image with a pure translation: out[r, c] = im(r - esy, c - esx)
        im = warp_bilinear(im, 1.0, -esy, -esx, im.shape)
    cx, cy = int(y_z) + res // 2, int(x_z) + res // 2
    sub = im[cx - n // 2:cx + n // 2, cy - n // 2:cy + n // 2]
    if np.isfinite(src_altitude):
        m = (src_altitude - layer_altitude) / src_altitude
        ratio = 1.0 / m
        extra = (ratio % 1) / 2 + (np.floor(ratio) - 1) * 0.5
        # scale about the image centre n/2 then translate by extra: out(p) = sub((p - n/2 - extra) / ratio + n/2)
        b = -(n / 2 + extra) / ratio + n / 2
        sub = warp_bilinear(sub, 1.0 / ratio, b, b, (n, n))
    return sub


def asm(field, wavelength, pitch, distance):
    """fftshift(ifft2(ifftshift(fftshift(fft2(ifftshift(E))) exp(-i pi lambda z f^2)))) (Atmosphere.py:938-972)."""
    if distance == 0:
        return field
    N = field.shape[0]
    f = np.arange(-N / 2, N / 2) / (N * pitch)
    fx, fy = np.meshgrid(f, f)
    H = np.exp(-1j * np.pi * wavelength * distance * (fx ** 2 + fy ** 2))
    F = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(field)))
    return np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(F * H)))


def scintillation_field(layer_opds, altitudes, wavelength, pitch):
    """Unit field multiplied by each layer's phase, highest layer first, propagated to the next layer and
    finally to the ground (Atmosphere.py:563-571, 644-662). layer_opds [m] and altitudes sorted top-down."""
    E = np.ones(layer_opds[0].shape, dtype=complex)
    for j, (opd, alt) in enumerate(zip(layer_opds, altitudes)):
        E = E * np.exp(1j * 2 * np.pi / wavelength * opd)
        dist = alt - altitudes[j + 1] if j < len(altitudes) - 1 else alt
        if dist > 1e-6:
            E = asm(E, wavelength, pitch, dist)
    return E

