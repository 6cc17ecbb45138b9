"""Pins of the Paganin oracle (oracle/paganin.py) -- CPU only."""
import numpy as np
import pytest

from oracle import paganin as opg
from oracle import propagation as op


def test_constant_intensity_closed_form():
    data = np.full((3, 12, 10), 0.8)
    ph = opg.multi_paganin(data, [1e-3, 2e-3, 3e-3], 1e-10, 1e-7, 50.0, 0.01)
    assert np.allclose(ph, 25.0 * np.log(0.8 / 1.01), rtol=1e-12, atol=1e-12)


def test_filter_equals_explicit_dft():
    rng = np.random.default_rng(0)
    ny, nx, nd = 6, 5, 2
    data = 1 + 0.1 * rng.standard_normal((nd, ny, nx))
    dist, lam, vs, db, al = [2e-3, 5e-3], 1e-10, 1e-7, 80.0, 0.02
    Fy = np.exp(-2j * np.pi * np.outer(np.arange(ny), np.arange(ny)) / ny)
    Fx = np.exp(-2j * np.pi * np.outer(np.arange(nx), np.arange(nx)) / nx)
    ky = np.array([k if k < (ny + 1) // 2 else k - ny for k in range(ny)]) / (ny * vs)
    kx = np.array([k if k < (nx + 1) // 2 else k - nx for k in range(nx)]) / (nx * vs)
    f2 = ky[:, None] ** 2 + kx[None, :] ** 2
    num = sum((1 + lam * z * np.pi * db * f2) * (Fy @ data[j] @ Fx.T) for j, z in enumerate(dist)) / nd
    den = sum((1 + lam * z * np.pi * db * f2) ** 2 for z in dist) / nd + al
    back = (np.conj(Fy) @ (num / den) @ np.conj(Fx).T) / (ny * nx)
    ref = np.log(np.real(back)) * db * 0.5
    assert np.allclose(opg.multi_paganin(data, dist, lam, vs, db, al), ref, rtol=1e-12, atol=1e-12)


def test_recovers_weak_phase_object_in_contact_regime():
    """Near-field intensities of exp(i phi - phi/db) (homogeneous object,
    beta = delta/db) at small Fresnel distances: Paganin returns phi up to
    the regularisation."""
    n = 64
    yy, xx = np.mgrid[0:n, 0:n] - n / 2
    phi = -0.05 * np.exp(-(xx ** 2 + yy ** 2) / (2 * 6.0 ** 2))
    db = 50.0
    psi = np.exp(1j * phi + phi / db)
    lam, vs = 1e-10, 1e-7
    dist = np.array([2e-3, 4e-3])
    prop = op.Propagation(n, n, 1, 2, lam, vs, dist)
    data = np.stack([np.abs(prop.D(psi[None], j)[0]) ** 2 for j in range(2)])
    ph = opg.multi_paganin(data, dist, lam, vs, db, 1e-8)
    assert np.linalg.norm(ph - phi) < 0.05 * np.linalg.norm(phi)
