"""Oracle: multi-distance Paganin phase retrieval (TEST INFRASTRUCTURE ONLY).

Restates experimental/AtomiumS2/steps15.py:599-615 (multiPaganin) in NumPy
float64, and the reflect padding / crop of its caller (steps15.py:715-745).

Pins (tests/test_oracle_paganin.py): a constant intensity c gives the closed
form (delta/beta)/2 log(c / (1 + alpha)); the FFT-domain filter equals an
explicit DFT-matrix evaluation of the same formula on a tiny grid; and a weak
pure-phase object propagated by the oracle's Fresnel operator is recovered in
the limit the Paganin formula is derived for (contact-regime TIE).
"""
import numpy as np


def multi_paganin(data, distances, wavelength, voxelsize, delta_beta, alpha):
    ny, nx = data.shape[-2:]
    fx = np.fft.fftfreq(nx, d=voxelsize)
    fy = np.fft.fftfreq(ny, d=voxelsize)
    fx, fy = np.meshgrid(fx, fy)
    num = 0
    den = 0
    for j in range(data.shape[0]):
        taylor = 1 + wavelength * distances[j] * np.pi * delta_beta * (fx ** 2 + fy ** 2)
        num = num + taylor * np.fft.fft2(data[j].astype(np.complex128))
        den = den + taylor ** 2
    num = num / len(distances)
    den = den / len(distances) + alpha
    return np.log(np.real(np.fft.ifft2(num / den))) * delta_beta * 0.5


def paganin_projection(data, distances, wavelength, voxelsize, delta_beta, alpha=0.01, pad=None):
    pad = data.shape[-1] // 8 if pad is None else pad
    padded = np.pad(data, ((0, 0), (pad, pad), (pad, pad)), "reflect") if pad else data
    ph = multi_paganin(padded, distances, wavelength, voxelsize, delta_beta, alpha)
    return ph[pad:pad + data.shape[-2], pad:pad + data.shape[-1]] if pad else ph
