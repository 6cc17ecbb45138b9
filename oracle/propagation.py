"""Oracle: Fresnel propagator D / DT (TEST INFRASTRUCTURE ONLY).

Restates src/holotomocupy/propagation.py:14-112 and the padding kernels
src/holotomocupy/cuda_kernels.py:78-127 step for step in NumPy:

  D(psi, j)  = crop(IFFT2_unnorm(FFT2(sympad(psi)) * fker[j]))
  DT(phi, j) = padadj(IFFT2_unnorm(FFT2(embed(phi)) * conj(fker[j])))

with fker[j] = exp(-i*pi*lambda*z_j*(fx^2+fy^2)) / (4*n*nz) on the 2nz x 2n
grid, fx = fftfreq(2n, d=voxelsize).

Pins (tests/test_oracle_propagation.py):
  * DCT-II eigenfunctions cos(pi*k*(x+1/2)/n) are eigenvectors of the 1-D
    factor with eigenvalue exp(-i*pi*lambda*z*(k/(2n*vs))^2) (closed form);
  * D(const) = const;
  * the 2-D operator equals the separable row-then-column composition;
  * DT equals D with the conjugated kernel and is its adjoint;
  * an independent DCT-domain implementation (scipy.fft.dct/idct, type 2,
    orthonormal) agrees to 1e-12;
  * pad_fwd equals numpy.pad(mode='symmetric') and pad_adj is its transpose.
"""
import numpy as np
import scipy.fft as sfft

WORKERS = -1


def fresnel_kernel(n, nz, wavelength, voxelsize, distance, ref32=False):
    """Transfer functions fker[ndist, 2nz, 2n]  (propagation.py:20-28).

    ref32=True emulates the reference's precision: fx, fy, f2 in float32 and
    the exponential in complex64 (propagation.py:20-28); otherwise float64.
    """
    distance = np.atleast_1d(np.asarray(distance, dtype=np.float64))
    fx = np.fft.fftfreq(2 * n, d=voxelsize)
    fy = np.fft.fftfreq(2 * nz, d=voxelsize)
    if ref32:
        fx = fx.astype(np.float32)
        fy = fy.astype(np.float32)
    fx, fy = np.meshgrid(fx, fy)
    f2 = fx ** 2 + fy ** 2
    norm = float(4 * n * nz)
    if ref32:
        out = np.empty([len(distance), 2 * nz, 2 * n], dtype=np.complex64)
        for j in range(len(distance)):
            arg = (-1j * np.pi * wavelength * distance[j]) * f2  # python complex * f32 -> c64
            out[j] = (np.exp(arg.astype(np.complex64)) / np.float32(norm)).astype(np.complex64)
        return out
    out = np.empty([len(distance), 2 * nz, 2 * n], dtype=np.complex128)
    for j in range(len(distance)):
        out[j] = np.exp(-1j * np.pi * wavelength * distance[j] * f2) / norm
    return out


def _sym_index(length_out, n):
    """Source index of the symmetric pad along one axis (cuda_kernels.py:89-92)."""
    t = np.arange(length_out)
    return np.where(t < n // 2, n // 2 - t - 1,
                    np.where(t >= n + n // 2, 2 * n - t + n // 2 - 1, t - n // 2))


def pad_fwd(f):
    """Symmetric pad [B, nz, n] -> [B, 2nz, 2n]  (cuda_kernels.py:78-98)."""
    _, nz, n = f.shape
    iy = _sym_index(2 * nz, nz)
    ix = _sym_index(2 * n, n)
    return f[:, iy[:, None], ix[None, :]]


def pad_adj(g):
    """Adjoint of pad_fwd: sum of the 4 mirror images  (cuda_kernels.py:100-127)."""
    B, nz2, n2 = g.shape
    nz, n = nz2 // 2, n2 // 2
    tx = np.arange(n)
    ty = np.arange(nz)
    gx_c = tx + n // 2
    gx_m = np.where(tx < n // 2, n // 2 - 1 - tx, 2 * n + n // 2 - 1 - tx)
    gy_c = ty + nz // 2
    gy_m = np.where(ty < nz // 2, nz // 2 - 1 - ty, 2 * nz + nz // 2 - 1 - ty)
    return (g[:, gy_c[:, None], gx_c[None, :]] + g[:, gy_c[:, None], gx_m[None, :]]
            + g[:, gy_m[:, None], gx_c[None, :]] + g[:, gy_m[:, None], gx_m[None, :]])


class Propagation:
    """Restatement of Propagation (propagation.py:11-112)."""

    def __init__(self, n, nz, ntheta, ndist, wavelength, voxelsize, distance,
                 dtype=np.complex128, ref32=False):
        self.n = n
        self.nz = nz
        self.ndist = ndist
        self.dtype = np.dtype(dtype)
        fk = fresnel_kernel(n, nz, wavelength, voxelsize, distance, ref32=ref32)
        self.fker = fk.astype(self.dtype)

    def _conv(self, buf, H):
        # cufft.fft2 -> *= fker -> ifft2(norm="forward")  (propagation.py:82-86)
        buf = sfft.fft2(buf, axes=(-2, -1), workers=WORKERS)
        buf *= H
        return sfft.ifft2(buf, axes=(-2, -1), norm="forward", workers=WORKERS)

    def D(self, psi, j):
        """Forward propagator  (propagation.py:70-89)."""
        added = psi.ndim == 2
        if added:
            psi = psi[None]
        psi = np.asarray(psi, dtype=self.dtype)
        buf = pad_fwd(psi)
        buf = self._conv(buf, self.fker[j])
        nz, n = self.nz, self.n
        res = np.ascontiguousarray(buf[:, nz // 2: -nz // 2, n // 2: -n // 2]).astype(self.dtype)
        return res[0] if added else res

    def DT(self, big_psi, j):
        """Adjoint propagator  (propagation.py:91-112)."""
        added = big_psi.ndim == 2
        if added:
            big_psi = big_psi[None]
        B = big_psi.shape[0]
        nz, n = self.nz, self.n
        buf = np.zeros([B, 2 * nz, 2 * n], dtype=self.dtype)
        buf[:, nz // 2: -nz // 2, n // 2: -n // 2] = big_psi
        buf = self._conv(buf, np.conj(self.fker[j]))
        res = pad_adj(buf).astype(self.dtype)
        return res[0] if added else res


def dct_domain_D(psi, wavelength, voxelsize, distance, adjoint=False):
    """Independent second oracle: D = (Q_y^T L_y Q_y) (x) (Q_x^T L_x Q_x).

    Q is the orthonormal DCT-II (scipy.fft.dct type 2) and L the diagonal
    exp(-+ i*pi*lambda*z*(k/(2n*vs))^2), k = 0..n-1.  Used only to pin the
    restatement above; it shares no code with it.
    """
    psi = np.asarray(psi, dtype=np.complex128)
    nz, n = psi.shape[-2:]
    s = 1.0 if adjoint else -1.0
    kx = np.arange(n) / (2 * n * voxelsize)
    ky = np.arange(nz) / (2 * nz * voxelsize)
    lx = np.exp(s * 1j * np.pi * wavelength * distance * kx ** 2)
    ly = np.exp(s * 1j * np.pi * wavelength * distance * ky ** 2)
    c = sfft.dct(psi, type=2, norm="ortho", axis=-1) * lx
    c = sfft.idct(c, type=2, norm="ortho", axis=-1)
    c = sfft.dct(c, type=2, norm="ortho", axis=-2) * ly[:, None]
    return sfft.idct(c, type=2, norm="ortho", axis=-2)
