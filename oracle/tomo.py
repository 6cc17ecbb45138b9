"""Oracle: USFFT Radon transform R / RT (TEST INFRASTRUCTURE ONLY).

Restates src/holotomocupy/tomo.py:12-116 (Tomo.__init__, R, RT) and the
gridding kernel src/holotomocupy/cuda_kernels.py:3-76 (gather, dir 0 and the
atomic scatter dir 1) in NumPy, float64 / complex128:

  R(obj)  = c1d IFFT1(c1d gather(c2d FFT2(c2d embed(obj phi))) / n) / 4
  RT(d)   = crop(c2d IFFT2(c2d scatter(c1d FFT1(c1d d)))) phi

with c1d[k] = (-1)^(k+1), c2d[i,j] = (-1)^(i+j), the embed placing the n x n
slice in the centre of a 2n x 2n grid, and 9x9 Gaussian taps
(pi/mu) exp(-(pi^2/mu) d^2) around each polar sample
x0 = (t - n/2) cos(theta)/n, y0 = -(t - n/2) sin(theta)/n.

Also restated: the detector-axis filter, filtered back-projection and CG
tomography of tomo.py:118-210 (filter_sino, fbp, rec_tomo).  The reference's
Parzen window takes u = 2 |f| 4n (not 2 |f|), so it is not the textbook Parzen
window; the restatement keeps the reference's formula.

Pins (tests/test_oracle_tomo.py): RT is the exact adjoint of R; an off-centre
Gaussian blob projects onto the trajectory t = n/2 + n (x_c cos - y_c sin)
(sign of cuda_kernels.py:22-23) and a centred one gives angle-independent
profiles; R of a smooth object equals its direct FP64 line integrals (in
pixels) divided by sqrt(n ntheta), to the USFFT accuracy (eps = 1e-3), which
pins the absolute normalisation; a brute-force FP64 type-2 NUDFT of the
deapodised, zero-padded slice at the polar points agrees with the gridded
spectrum to the accuracy the stated eps / mu / m give, and a changed mu fails
it; the filters act on cos(2 pi k t / n) with the closed-form eigenvalues
4|k|, 4|k| sinc(k/n) and 4|k| 2 (1 - 8|k|)^3; FBP of R of a smooth blob
returns the blob; rec_tomo's first step is the exact line minimiser along the
gradient and its search directions are R-conjugate.
"""
import math

import numpy as np
import scipy.fft as sfft

WORKERS = -1


def usfft_params(n):
    """eps, mu, m (tomo.py:14-16)."""
    eps = 1e-3
    mu = -math.log(eps) / (2 * n * n)
    m = math.ceil(2 * n / math.pi * math.sqrt(-mu * math.log(eps) + (mu * n) ** 2 / 4))
    return eps, mu, m


def smooth_mask(n, mask_r):
    """Support mask of tomo.py:34-47 (FP64 NumPy, cast to float32 as there)."""
    if not mask_r > 0:
        return np.ones((n, n))
    t1d = np.linspace(-1, 1, n)
    x, y = np.meshgrid(t1d, t1d)
    circ = (x ** 2 + y ** 2 < mask_r).astype("float32")
    g = np.exp(-(20 ** 2) * (x ** 2 + y ** 2))
    fcirc = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(circ)))
    fg = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(g)))
    mask = np.fft.fftshift(np.fft.ifft2(np.fft.fftshift(fcirc * fg))).real.astype("float32")
    mask /= np.amax(mask)
    return mask.astype(np.float64)


def deapodizer(n, ntheta, mask_r):
    """phi of tomo.py:19-21, 46-47."""
    _, mu, _ = usfft_params(n)
    t = np.linspace(-1 / 2, 1 / 2, n, endpoint=False)
    dx, dy = np.meshgrid(t, t)
    phi = np.exp(mu * (n * n) * (dx * dx + dy * dy)) * (1 - n % 4)
    return phi * smooth_mask(n, mask_r) / (n * np.sqrt(n * ntheta))


def signs(n):
    c1d = 1 - 2 * (np.arange(1, n + 1) % 2)
    c2 = 1 - 2 * (np.arange(1, 2 * n + 1) % 2)
    return c1d.astype(np.float64), np.outer(c2, c2).astype(np.float64)


def _taps(n, theta):
    """Per (theta, t): grid rows/cols and separable weights of the 81 taps
    (cuda_kernels.py:17-51).  Returns ry [T, n, L], cx [T, n, L], wy, wx."""
    _, mu, m = usfft_params(n)
    L = 2 * m + 1
    tt = np.arange(n)
    x0 = (tt[None, :] - n // 2) / n * np.cos(theta)[:, None]
    y0 = -(tt[None, :] - n // 2) / n * np.sin(theta)[:, None]
    bx = np.floor(2 * n * x0).astype(np.int64) - m
    by = np.floor(2 * n * y0).astype(np.int64) - m
    k = np.arange(L)
    ell0 = bx[..., None] + k
    ell1 = by[..., None] + k
    wx = np.exp(-math.pi ** 2 / mu * (ell0 / (2 * n) - x0[..., None]) ** 2)
    wy = (math.pi / mu) * np.exp(-math.pi ** 2 / mu * (ell1 / (2 * n) - y0[..., None]) ** 2)
    cx = (n + ell0 + 2 * n) % (2 * n)
    ry = (n + ell1 + 2 * n) % (2 * n)
    return ry, cx, wy, wx


class Tomo:
    """Restatement of Tomo (tomo.py:12-116) in FP64."""

    def __init__(self, n, nz, theta, mask_r):
        self.n = n
        self.nz = nz
        # the reference keeps the angles as float32 (tomo.py:32); the taps are
        # then evaluated in FP64 at those angles
        self.theta = np.asarray(theta, dtype=np.float64).astype(np.float32).astype(np.float64)
        self.ntheta = len(self.theta)
        self.phi = deapodizer(n, self.ntheta, mask_r)
        self.c1d, self.c2d = signs(n)
        self.taps = _taps(n, self.theta)

    def _gather(self, G):
        """dir 0: Shat[theta, z, t] = sum w G[z, row, col] / n."""
        ry, cx, wy, wx = self.taps
        # G[z, ry[...,a], cx[...,b]] * wy[a] wx[b]
        vals = G[:, ry[:, :, :, None], cx[:, :, None, :]]  # [z, T, n, L, L]
        s = np.einsum("ztnab,tna,tnb->tzn", vals, wy, wx, optimize=True)
        return s / self.n

    def _scatter(self, S, nz):
        """dir 1: G[z, row, col] += w Shat[theta, z, t] (atomic scatter)."""
        ry, cx, wy, wx = self.taps
        n = self.n
        G = np.zeros((nz, 2 * n, 2 * n), dtype=np.complex128)
        L = wy.shape[-1]
        w = wy[:, :, :, None] * wx[:, :, None, :]  # [T, n, L, L]
        rows = np.broadcast_to(ry[:, :, :, None], w.shape).reshape(-1)
        cols = np.broadcast_to(cx[:, :, None, :], w.shape).reshape(-1)
        flat = rows * (2 * n) + cols
        for z in range(nz):
            contrib = (w * S[:, z, :, None, None]).reshape(-1)
            g = np.zeros(4 * n * n, dtype=np.complex128)
            np.add.at(g, flat, contrib)
            G[z] = g.reshape(2 * n, 2 * n)
        del L
        return G

    def R(self, obj):
        """tomo.py:56-86."""
        obj = np.asarray(obj)
        nz, n = obj.shape[0], self.n
        fde = np.zeros((nz, 2 * n, 2 * n), dtype=np.complex128)
        fde[:, n // 2: 3 * n // 2, n // 2: 3 * n // 2] = obj * self.phi
        fde = self.c2d * sfft.fft2(self.c2d * fde, axes=(-2, -1), workers=WORKERS)
        sino = self._gather(fde)
        sino = self.c1d * sfft.ifft(self.c1d * sino, axis=-1, workers=WORKERS)
        res = sino / 4
        return res.real if np.isrealobj(obj) else res

    def RT(self, data):
        """tomo.py:88-116."""
        data = np.asarray(data)
        nz, n = data.shape[1], self.n
        sino = self.c1d * sfft.fft(self.c1d * data.astype(np.complex128), axis=-1, workers=WORKERS)
        fde = self._scatter(sino, nz)
        fde = self.c2d * sfft.ifft2(self.c2d * fde, axes=(-2, -1), workers=WORKERS)
        res = fde[:, n // 2: 3 * n // 2, n // 2: 3 * n // 2] * self.phi
        return res.real if np.isrealobj(data) else res

    # ---- filtered back-projection and CG tomography (tomo.py:118-210)
    def filter_h(self, filter_name):
        """Frequency response h[k] on fftfreq(n) (tomo.py:131-155): ramp 4n|f|,
        Shepp-Logan 4n|f| sinc(f), Parzen 4n|f| w(u) with u = 2 (4n|f|) and
        w(u) = 1 - 6u^2 + 6u^3 (u <= 1/2), 2 (1 - u)^3 otherwise."""
        n = self.n
        f = np.fft.fftfreq(n)
        af = np.abs(f) * 4 * n
        if filter_name == "ramp":
            return af
        if filter_name == "shepp":
            return af * np.sinc(f)
        if filter_name == "parzen":
            u = 2 * af
            return af * np.where(u <= 0.5, 1 - 6 * u ** 2 + 6 * u ** 3, 2 * (1 - u) ** 3)
        raise ValueError(f"Unknown filter '{filter_name}'. Choose: ramp, shepp, parzen.")

    def filter_sino(self, data, filter_name):
        """_filter_sino (tomo.py:118-162): ifft(fft(data) h) along the detector
        axis; the real part for a real sinogram."""
        data = np.asarray(data)
        h = self.filter_h(filter_name)
        res = sfft.ifft(sfft.fft(data.astype(np.complex128), axis=-1, workers=WORKERS) * h, axis=-1,
                        workers=WORKERS)
        return res.real if np.isrealobj(data) else res

    def fbp(self, data, filter_name="ramp"):
        """fbp (tomo.py:164-186): RT(filter_sino(data)) sqrt(n / ntheta)."""
        return self.RT(self.filter_sino(data, filter_name)) * np.sqrt(self.n / self.ntheta)

    def rec_tomo(self, d, niter=1, trace=None):
        """rec_tomo (tomo.py:188-210): conjugate gradients on |R u - d|^2 from
        u = 0, beta = Re<R grad, R eta> / Re<R eta, R eta>,
        alpha = -Re<grad, eta> / (2 Re<R eta, R eta>).  trace (a list) receives
        (u, grad, eta, Reta, alpha) of every iteration."""
        d = np.asarray(d)
        dt = np.float64 if np.isrealobj(d) else np.complex128
        u = np.zeros((d.shape[1], self.n, self.n), dtype=dt)
        Ru = self.R(u)
        eta = Reta = None
        for k in range(niter):
            grad = self.RT(2 * (Ru - d))
            Rgrad = self.R(grad)
            if k == 0:
                eta, Reta = -grad, -Rgrad
            else:
                beta = _redot(Rgrad, Reta) / _redot(Reta, Reta)
                eta = beta * eta - grad
                Reta = beta * Reta - Rgrad
            alpha = -_redot(grad, eta) / (2 * _redot(Reta, Reta))
            u = u + alpha * eta
            Ru = Ru + alpha * Reta
            if trace is not None:
                trace.append((u.copy(), grad, eta.copy(), Reta.copy(), alpha))
        return u


def _redot(a, b):
    """Re<a, b> (utils.py:96-103)."""
    return float(np.sum(np.real(np.conj(a) * b)))
