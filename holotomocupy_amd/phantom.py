"""Seeded synthetic inputs for the BH operator stack (host NumPy; not on the
hot path): probe, projected-object fields, positions and measured amplitudes
of the holotomography forward model (the quantities rec_mpi.Rec consumes,
src/holotomocupy/rec_mpi.py:105-136)."""
import numpy as np


def ellipses(rng, shape, count=20, amp=(0.0, 1.0)):
    """Sum of `count` random filled ellipses on a [nz, n] grid: centres
    U[-0.3, 0.3]*size, semi-axes U[0.03, 0.2]*size, angle U[0, pi),
    amplitude U[amp]."""
    nz, n = shape
    yy, xx = np.mgrid[0:nz, 0:n].astype(np.float64)
    yy -= nz / 2
    xx -= n / 2
    out = np.zeros(shape)
    for _ in range(count):
        cy, cx = rng.uniform(-0.3, 0.3, 2) * np.array([nz, n])
        ay, ax = rng.uniform(0.03, 0.2, 2) * np.array([nz, n])
        t = rng.uniform(0, np.pi)
        u = (xx - cx) * np.cos(t) + (yy - cy) * np.sin(t)
        v = -(xx - cx) * np.sin(t) + (yy - cy) * np.cos(t)
        out += rng.uniform(*amp) * ((u / ax) ** 2 + (v / ay) ** 2 <= 1)
    return out


def smooth_noise(rng, shape, sigma):
    """Gaussian-smoothed white noise (spectral filter), unit RMS."""
    nz, n = shape
    f = np.fft.fft2(rng.standard_normal(shape))
    ky = np.fft.fftfreq(nz)[:, None]
    kx = np.fft.fftfreq(n)[None, :]
    f *= np.exp(-2 * (np.pi * sigma) ** 2 * (kx ** 2 + ky ** 2))
    v = np.fft.ifft2(f).real
    return v / (v.std() + 1e-30)


def probe(rng, ndist, nz, n, width=0.3, phase_rms=0.5, sigma=16.0):
    """Gaussian amplitude exp(-r^2/(2 (width*n)^2)) times exp(i phi), phi a
    smoothed random phase of the given RMS; one probe per distance."""
    yy, xx = np.mgrid[0:nz, 0:n].astype(np.float64)
    r2 = (xx - n / 2) ** 2 + (yy - nz / 2) ** 2
    amp = np.exp(-r2 / (2 * (width * n) ** 2))
    out = np.empty((ndist, nz, n), np.complex64)
    for j in range(ndist):
        out[j] = amp * np.exp(1j * phase_rms * smooth_noise(rng, (nz, n), sigma))
    return out


def projections(rng, B, nzobj, nobj, count=12, beta_ratio=0.01, dtype="complex64"):
    """Projected object fields [B, nzobj, nobj]: phase from random ellipses
    (U[0, 1] rad each), absorption beta_ratio * phase; real phase only for a
    float32 object."""
    base = ellipses(rng, (nzobj, nobj), count)
    out = np.empty((B, nzobj, nobj), np.complex64 if dtype == "complex64" else np.float32)
    for b in range(B):
        ph = np.roll(base, (b * 3) % nzobj, axis=1) + 0.1 * smooth_noise(rng, (nzobj, nobj), 4.0)
        out[b] = ph + 1j * beta_ratio * ph if dtype == "complex64" else ph
    return out


def positions(rng, B, ndist, amp=3.0):
    """Per-angle, per-distance shifts U[-amp, amp] px, (y, x)."""
    return rng.uniform(-amp, amp, (B, ndist, 2)).astype(np.float32)
