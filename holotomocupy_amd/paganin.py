"""Multi-distance Paganin phase retrieval, the initial guess of the BH solver.

Restates multiPaganin of the reference's data preparation
(experimental/AtomiumS2/steps15.py:599-615) on the device:

  phase = (delta/beta) / 2 * log(Re IFFT2( mean_j[T_j F(I_j)] / (mean_j[T_j^2] + alpha) )),
  T_j = 1 + lambda z_j pi (delta/beta) (fx^2 + fy^2),  fx = fftfreq(nx, voxelsize)

and its use there (steps15.py:715-745): reflect padding by nobj/8 on each side
before the filter and the crop after it.  An initialisation step that runs
once per projection, outside the BH iteration: the 2-D FFTs of the padded
[ndist, 1.25 n, 1.25 n] stacks (not a power of two) run on the device FFT.
"""
import numpy as np
import torch

from . import _lib


def multi_paganin(data, distances, wavelength, voxelsize, delta_beta, alpha):
    """data [ndist, ny, nx] float32 intensities (device tensor) -> phase
    [ny, nx] float32 (steps15.py:599-615)."""
    _lib.require_gpu()
    if data.ndim != 3 or data.shape[0] != len(distances):
        raise ValueError("data must be [ndist, ny, nx] with one distance per plane")
    dev = data.device
    ny, nx = data.shape[-2:]
    fx = torch.fft.fftfreq(nx, d=voxelsize, device=dev).float()
    fy = torch.fft.fftfreq(ny, d=voxelsize, device=dev).float()
    fy, fx = torch.meshgrid(fy, fx, indexing="ij")
    f2 = fx ** 2 + fy ** 2
    num = torch.zeros((ny, nx), dtype=torch.complex64, device=dev)
    den = torch.zeros((ny, nx), dtype=torch.float32, device=dev)
    for j in range(data.shape[0]):
        taylor = 1 + float(wavelength * distances[j] * np.pi * delta_beta) * f2
        num += taylor * torch.fft.fft2(data[j].to(torch.complex64))
        den += taylor ** 2
    num /= len(distances)
    den = den / len(distances) + alpha
    phase = torch.log(torch.real(torch.fft.ifft2(num / den)))
    return phase * (delta_beta * 0.5)


def paganin_projection(data, distances, wavelength, voxelsize, delta_beta, alpha=0.01, pad=None):
    """Reflect-pad data [ndist, ny, nx] by pad (default nx // 8) on each side,
    retrieve the phase and crop back to [ny, nx] (steps15.py:717-745)."""
    pad = data.shape[-1] // 8 if pad is None else pad
    padded = torch.nn.functional.pad(data[None], (pad, pad, pad, pad), mode="reflect")[0] if pad else data
    ph = multi_paganin(padded, distances, wavelength, voxelsize, delta_beta, alpha)
    return ph[pad:pad + data.shape[-2], pad:pad + data.shape[-1]] if pad else ph
