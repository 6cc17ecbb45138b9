"""Oracle: cubic B-spline shift family (TEST INFRASTRUCTURE ONLY).

Restates src/holotomocupy/shift.py:18-221 (with the back-shift Sback /
coeffback of shift.py:84-153 and its sback kernel, cuda_kernels.py:294-368)
and the device kernels
src/holotomocupy/cuda_kernels.py:131-290 (s / sf), 377-542 (d2s / d2sf),
547-687 (ds / dsf) and 690-843 (dsadj / dsadjf).

Every kernel of the reference evaluates, per output pixel (tx, ty) of angle b,
the sampling point (cuda_kernels.py:172-181)
    x = m_b (tx - n//2)  - r_b[1] + (m_b - 1)/2 + npsi//2
    y = m_b (ty - nz//2) - r_b[0] + (m_b - 1)/2 + nzpsi//2
and the 4x4 taps (floor(y)+jy, floor(x)+jx), jy, jx in {-1,0,1,2}, weighted by
products phi(dy - jy) * phi(dx - jx) (or derivatives), skipping taps outside
the grid.  Because every weight is such a product, each operator is a sum of
terms  Wy_b  c_b  Wx_b^T  with sparse per-axis matrices (4 entries per row);
this module builds those matrices and applies them, which is the same sum in
another order.  The adjoints use the transposed matrices -- the reference's
atomic scatter (dir = 1) computes exactly Wy^T g Wx.

Pins (tests/test_oracle_shift.py; the back-shift in tests/test_oracle_sback.py):
  * S(coeff(psi), r, 1) equals scipy.ndimage.shift(psi, -r, order=3,
    mode='grid-wrap') on compactly supported inputs;
  * S(c, r, m) equals scipy.ndimage.map_coordinates(36 c, ..., order=3,
    prefilter=False, mode='grid-constant') with magnification;
  * integer shifts at m = 1 reproduce shifted pixels exactly;
  * <S c, g> = <c, Sadj g>; coeff is self-adjoint and inverts the periodic
    [1 4 1] circulant;
  * dS, d2S match central finite differences in (c, r) (Taylor ratios 4 / 8);
  * dSadj satisfies <dS(c1, dr), g> = <c1, out1> + dr . out2.
"""
import numpy as np
import scipy.fft as sfft
import scipy.sparse as sp

WORKERS = -1


def phi(t):
    """6 * cubic B-spline  (cuda_kernels.py:131-139, shift.py:54-59)."""
    t = np.asarray(t, dtype=np.float64)
    a = np.abs(t)
    return np.where((-2 < t) & (t <= -1), (t + 2) ** 3,
                    np.where((-1 < t) & (t <= 1), 4 - 6 * t * t + 3 * a * t * t,
                             np.where((1 < t) & (t <= 2), (2 - t) ** 3, 0.0)))


def dphi(t):
    """(cuda_kernels.py:141-149)"""
    t = np.asarray(t, dtype=np.float64)
    a = np.abs(t)
    return np.where((-2 < t) & (t <= -1), 3 * (t + 2) ** 2,
                    np.where((-1 < t) & (t <= 1), -12 * t + 9 * a * t,
                             np.where((1 < t) & (t <= 2), -3 * (2 - t) ** 2, 0.0)))


def d2phi(t):
    """(cuda_kernels.py:151-159)"""
    t = np.asarray(t, dtype=np.float64)
    a = np.abs(t)
    return np.where((-2 < t) & (t <= -1), 6 * (t + 2),
                    np.where((-1 < t) & (t <= 1), -12 + 18 * a,
                             np.where((1 < t) & (t <= 2), 6 * (2 - t), 0.0)))


def ifB3(nzpsi, npsi):
    """Periodic prefilter response  (shift.py:25-30)."""
    x = np.linspace(-1 / 2, 1 / 2 - 1 / npsi, npsi)
    y = np.linspace(-1 / 2, 1 / 2 - 1 / nzpsi, nzpsi)
    divx = phi(0) + 2 * phi(1) * np.cos(2 * np.pi * x)
    divy = phi(0) + 2 * phi(1) * np.cos(2 * np.pi * y)
    return 1 / np.fft.fftshift(np.outer(divy, divx), axes=(-1, -2))


def coeff(psi, real=False, single=False):
    """B-spline coefficients: ifft2(fft2(psi) * ifB3)  (shift.py:74-82).
    single: complex64 transforms (the reference's precision), else complex128."""
    psi = np.asarray(psi)
    nzpsi, npsi = psi.shape[-2:]
    h = ifB3(nzpsi, npsi)
    if single:
        psi, h = psi.astype(np.complex64), h.astype(np.float32)
    out = sfft.ifft2(sfft.fft2(psi, axes=(-2, -1), workers=WORKERS) * h, axes=(-2, -1), workers=WORKERS)
    return out.real if real else out


def axis_matrix(npix, npsi, coord, kind=0):
    """Sparse [npix, npsi] matrix of one axis: row t holds the 4 weights
    phi^(kind)(d - j) at columns floor(coord[t]) + j, j = -1..2, in range."""
    f = (phi, dphi, d2phi)[kind]
    i0 = np.floor(coord)
    d = coord - i0
    rows, cols, vals = [], [], []
    t = np.arange(npix)
    for j in range(-1, 3):
        idx = (i0 + j).astype(np.int64)
        ok = (idx >= 0) & (idx < npsi)
        rows.append(t[ok])
        cols.append(idx[ok])
        vals.append(f(d[ok] - j))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(npix, npsi))


def coords(n, nz, npsi, nzpsi, r, m):
    """Per-angle sampling coordinates x[B, n], y[B, nz]  (cuda_kernels.py:172-181)."""
    r = np.asarray(r, dtype=np.float64).reshape(-1, 2)
    m = np.asarray(m, dtype=np.float64).reshape(-1)
    half = (m - 1) / 2
    x = m[:, None] * (np.arange(n) - n // 2)[None] - r[:, 1:2] + half[:, None] + npsi // 2
    y = m[:, None] * (np.arange(nz) - nz // 2)[None] - r[:, 0:1] + half[:, None] + nzpsi // 2
    return x, y


class Shift:
    """Restatement of Shift (shift.py:15-221) on float64 / complex128, or with
    single=True on float32 / complex64 (FP32 weights and sums, for timing the
    reference's precision on the CPU; the sampling coordinates stay FP64)."""

    def __init__(self, n, npsi, nz, nzpsi, obj_dtype="complex64", nchunk=None, single=False):
        self.n, self.npsi, self.nz, self.nzpsi = n, npsi, nz, nzpsi
        self.real = obj_dtype == "float32"
        self.single = single
        if single:
            self.dtype = np.float32 if self.real else np.complex64
        else:
            self.dtype = np.float64 if self.real else np.complex128

    def coeff(self, psi):
        return coeff(psi, real=self.real, single=self.single)

    def _mats(self, r, m, kinds=(0,)):
        x, y = coords(self.n, self.nz, self.npsi, self.nzpsi, r, m)
        B = x.shape[0]
        wt = np.float32 if self.single else np.float64
        wx = [[axis_matrix(self.n, self.npsi, x[b], k).astype(wt) for k in kinds] for b in range(B)]
        wy = [[axis_matrix(self.nz, self.nzpsi, y[b], k).astype(wt) for k in kinds] for b in range(B)]
        return wx, wy

    @staticmethod
    def _fwd(Wy, c, Wx):
        # Wy c Wx^T  for one angle
        return np.asarray((Wx @ (Wy @ c).T).T)

    @staticmethod
    def _adj(Wy, g, Wx):
        # Wy^T g Wx
        return np.asarray((Wx.T @ (Wy.T @ g).T).T)

    def S(self, c, r, m):
        """cuda_kernels.py:161-290 (dir = 0); shift.py:102-111."""
        c = np.asarray(c)
        wx, wy = self._mats(r, m)
        out = np.stack([self._fwd(wy[b][0], c[b], wx[b][0]) for b in range(c.shape[0])])
        return out.astype(self.dtype)

    curlySc = S

    def Sadj(self, g, r, m):
        """cuda_kernels.py:161-290 (dir = 1, atomic scatter); shift.py:113-122."""
        g = np.asarray(g)
        wx, wy = self._mats(r, m)
        out = np.stack([self._adj(wy[b][0], g[b], wx[b][0]) for b in range(g.shape[0])])
        return out.astype(self.dtype)

    def curlyS(self, psi, r, m):
        """shift.py:128-132."""
        return self.S(self.coeff(psi), r, m)

    def dcurlySc(self, c, r, m, c1, Deltar):
        """First differential  (ds kernel, cuda_kernels.py:547-620; shift.py:170-183):
        sum[-(phi'_x phi_y drx + phi_x phi'_y dry) c + phi_x phi_y c1]."""
        wx, wy = self._mats(r, m, kinds=(0, 1))
        dr = np.asarray(Deltar, dtype=np.float64).reshape(-1, 2)
        out = []
        for b in range(len(c)):
            (Px, Dx), (Py, Dy) = wx[b], wy[b]
            v = (-dr[b, 1] * self._fwd(Py, c[b], Dx) - dr[b, 0] * self._fwd(Dy, c[b], Px)
                 + self._fwd(Py, c1[b], Px))
            out.append(v)
        return np.stack(out).astype(self.dtype)

    def d2curlySc(self, c, r, m, c1, Deltar1, c2, Deltar2):
        """Second-order term  (d2s kernel, cuda_kernels.py:377-455; shift.py:206-221)."""
        wx, wy = self._mats(r, m, kinds=(0, 1, 2))
        d1 = np.asarray(Deltar1, dtype=np.float64).reshape(-1, 2)
        d2 = np.asarray(Deltar2, dtype=np.float64).reshape(-1, 2)
        out = []
        for b in range(len(c)):
            (Px, Dx, DDx), (Py, Dy, DDy) = wx[b], wy[b]
            d1x, d1y, d2x, d2y = d1[b, 1], d1[b, 0], d2[b, 1], d2[b, 0]
            cross = d1x * d2y + d1y * d2x
            v = (d1x * d2x * self._fwd(Py, c[b], DDx) + cross * self._fwd(Dy, c[b], Dx)
                 + d1y * d2y * self._fwd(DDy, c[b], Px))
            v -= d1x * self._fwd(Py, c1[b], Dx) + d1y * self._fwd(Dy, c1[b], Px)
            v -= d2x * self._fwd(Py, c2[b], Dx) + d2y * self._fwd(Dy, c2[b], Px)
            out.append(v)
        return np.stack(out).astype(self.dtype)

    def dcurlySadjc(self, c, r, m, Deltaphi):
        """Adjoint of the differential  (dsadj kernel, cuda_kernels.py:690-766;
        shift.py:185-204): [Sadj(Deltaphi), out2] with
        out2[:, 0] = Re<Deltaphi, -sum phi_x phi'_y c>,
        out2[:, 1] = Re<Deltaphi, -sum phi'_x phi_y c>."""
        wx, wy = self._mats(r, m, kinds=(0, 1))
        out1, out2 = [], np.zeros((len(c), 2))
        for b in range(len(c)):
            (Px, Dx), (Py, Dy) = wx[b], wy[b]
            g = Deltaphi[b]
            out1.append(self._adj(Py, g, Px))
            dt1 = -self._fwd(Dy, c[b], Px)
            dt2 = -self._fwd(Py, c[b], Dx)
            out2[b, 0] = np.sum(np.real(g) * np.real(dt1) + np.imag(g) * np.imag(dt1))
            out2[b, 1] = np.sum(np.real(g) * np.real(dt2) + np.imag(g) * np.imag(dt2))
        return [np.stack(out1).astype(self.dtype), out2]


# ---------------------------------------------------------------- back-shift (initial guess)
def phi_back(t, m):
    """phi(t / m) of the sback kernel (cuda_kernels.py:297-304)."""
    return phi(np.asarray(t, dtype=np.float64) / m)


def back_axis_matrix(npix, npsi, coord, m):
    """Sparse [npix, npsi] matrix of one axis of the sback kernel
    (cuda_kernels.py:316-356): row t holds phi((d - j)/m) at columns
    floor(coord[t]) + j for j in [ceil(d - 2m), (int)(d + 2m)) (the upper
    bound exclusive, as the kernel's loop has it), in range."""
    i0 = np.floor(coord)
    d = coord - i0
    span = 2.0 * m
    rows, cols, vals = [], [], []
    for t in range(npix):
        for j in range(int(np.ceil(d[t] - span)), int(np.floor(d[t] + span))):
            idx = int(i0[t]) + j
            if 0 <= idx < npsi:
                rows.append(t)
                cols.append(idx)
                vals.append(phi_back(d[t] - j, m))
    return sp.csr_matrix((np.asarray(vals, dtype=np.float64), (rows, cols)), shape=(npix, npsi))


def sback_mats(n, npsi, nz, nzpsi, r, m):
    x, y = coords(n, nz, npsi, nzpsi, r, m)
    m = np.broadcast_to(np.asarray(m, dtype=np.float64).reshape(-1), (x.shape[0],))
    return ([back_axis_matrix(nz, nzpsi, y[b], m[b]) for b in range(x.shape[0])],
            [back_axis_matrix(n, npsi, x[b], m[b]) for b in range(x.shape[0])])


def sback(g, r, m, n, npsi, nz, nzpsi):
    """Shift.Sback (shift.py:138-149): the sback scatter (dir 1) as Wy^T g Wx."""
    g = np.asarray(g, dtype=np.complex128)
    Wy, Wx = sback_mats(n, npsi, nz, nzpsi, r, m)
    return np.stack([Wy[b].T @ g[b] @ Wx[b] for b in range(g.shape[0])])


def sback_forward(c, r, m, n, npsi, nz, nzpsi):
    """dir 0 of the sback kernel (the gather whose transpose Sback is)."""
    c = np.asarray(c, dtype=np.complex128)
    Wy, Wx = sback_mats(n, npsi, nz, nzpsi, r, m)
    return np.stack([Wy[b] @ c[b] @ Wx[b].T for b in range(c.shape[0])])


def coeffback_filter(N, m):
    """D(f) = phi(0) + 2 sum_{k=1..4} phi(k/m) cos(2 pi k x) on
    x = linspace(-1/2, 1/2 - 1/N, N), fftshifted (shift.py:84-96)."""
    x = np.linspace(-0.5, 0.5 - 1.0 / N, N)
    d = phi(0.0) + sum(2 * phi(k / m) * np.cos(2 * np.pi * k * x) for k in range(1, 5))
    return np.fft.fftshift(d)


def coeffback(psi, m):
    """Shift.coeffback: IFFT2(FFT2(psi) / (D_y (x) D_x))."""
    psi = np.asarray(psi, dtype=np.complex128)
    H = 1.0 / np.outer(coeffback_filter(psi.shape[-2], m), coeffback_filter(psi.shape[-1], m))
    return sfft.ifft2(sfft.fft2(psi, axes=(-2, -1), workers=WORKERS) * H, axes=(-2, -1), workers=WORKERS)
