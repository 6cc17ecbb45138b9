"""Pins of the USFFT Radon oracle (oracle/tomo.py) -- CPU only."""
import numpy as np
import pytest
import scipy.ndimage as ndi

from oracle import tomo as ot


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_usfft_half_width_is_4():
    for n in (64, 128, 256, 1024, 4096):
        assert ot.usfft_params(n)[2] == 4


@pytest.mark.parametrize("mask_r", [0, 0.9])
def test_RT_is_adjoint_of_R(mask_r):
    rng = np.random.default_rng(0)
    n, nz = 32, 2
    theta = np.arange(24) * np.pi / 24
    T = ot.Tomo(n, nz, theta, mask_r)
    x = crandn(rng, (nz, n, n))
    y = crandn(rng, (24, nz, n))
    lhs = np.vdot(y, T.R(x))
    rhs = np.vdot(T.RT(y), x)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    xr = rng.standard_normal((nz, n, n))
    yr = rng.standard_normal((24, nz, n))
    assert abs(np.sum(T.R(xr) * yr) - np.sum(xr * T.RT(yr))) < 1e-12 * abs(np.sum(T.R(xr) * yr))


def blob(n, ic, jc, w):
    i, j = np.mgrid[0:n, 0:n]
    return np.exp(-((i - ic) ** 2 + (j - jc) ** 2) / (2 * w * w))


def test_blob_trajectory_and_centre_invariance():
    n = 64
    theta = np.arange(48) * np.pi / 48
    T = ot.Tomo(n, 1, theta, 0)
    ic, jc = n / 2 + 9.0, n / 2 - 13.0
    p = T.R(blob(n, ic, jc, 2.0)[None])[:, 0]
    t_peak = np.array([np.sum(np.arange(n) * np.abs(r) ** 4) / np.sum(np.abs(r) ** 4) for r in p])
    t_ref = n / 2 + (jc - n / 2) * np.cos(theta) - (ic - n / 2) * np.sin(theta)
    assert np.abs(t_peak - t_ref).max() < 0.2
    c = T.R(blob(n, n / 2, n / 2, 3.0)[None])[:, 0]
    assert np.abs(c - c[0]).max() < 2e-3 * np.abs(c).max()


@pytest.mark.parametrize("n,ntheta", [(64, 16), (64, 32), (128, 16)])
def test_matches_line_integrals_with_absolute_scale(n, ntheta):
    """Direct FP64 line integrals p(theta, s) = int f(x, y) along
    x cos - y sin = s (in pixels), against R: R = p / sqrt(n ntheta), the
    normalisation fixed by phi's 1/(n sqrt(n ntheta)) and the 1/4, 1/n of the
    gather and line transform (a dropped or doubled factor fails)."""
    theta = np.arange(ntheta) * np.pi / ntheta
    T = ot.Tomo(n, 1, theta, 0)
    f = blob(n, n / 2 + 5, n / 2 - 3, 4.0) + 0.5 * blob(n, n / 2 - 8, n / 2 + 6, 3.0)
    p = T.R(f[None])[:, 0].real
    s = np.arange(n) - n // 2
    u = np.linspace(-n, n, 4 * n + 1)
    ref = np.empty_like(p)
    for a, th in enumerate(theta):
        # points s (cos, -sin) + u (sin, cos) in (x, y); map to (row, col) = (y + n/2, x + n/2)
        xs = s[:, None] * np.cos(th) + u[None, :] * np.sin(th)
        ys = -s[:, None] * np.sin(th) + u[None, :] * np.cos(th)
        vals = ndi.map_coordinates(f, [ys + n / 2, xs + n / 2], order=3, mode="constant")
        ref[a] = np.trapezoid(vals, u, axis=1)
    ref /= np.sqrt(n * ntheta)
    assert np.linalg.norm(p - ref) / np.linalg.norm(p) < 5e-3
    scale = np.sum(p * ref) / np.sum(ref * ref)
    assert abs(scale - 1) < 1e-4
