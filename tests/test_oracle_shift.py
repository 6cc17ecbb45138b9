"""Pins of the B-spline oracle (oracle/shift.py) against scipy.ndimage,
exact identities and finite differences -- CPU only."""
import numpy as np
import pytest
import scipy.ndimage as ndi

from oracle import shift as osh


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def blob(n, rng):
    """Smooth complex field with compact support well inside the grid."""
    y, x = np.mgrid[0:n, 0:n] - n / 2
    r2 = (x ** 2 + y ** 2) / (0.18 * n) ** 2
    w = np.where(r2 < 1, (1 - r2) ** 4, 0.0)
    return w * (1.0 + 0.3 * np.cos(0.4 * x) + 0.5j * np.sin(0.3 * y + 0.2 * x))


def test_S_coeff_equals_ndimage_shift():
    rng = np.random.default_rng(0)
    n = 48
    sh = osh.Shift(n, n, n, n)
    psi = np.stack([blob(n, rng), blob(n, rng) * 0.5])
    r = np.array([[1.3, -2.7], [-0.45, 3.25]])
    out = sh.curlyS(psi, r, np.ones(2))
    for b in range(2):
        ref = (ndi.shift(psi[b].real, r[b], order=3, mode="grid-wrap")
               + 1j * ndi.shift(psi[b].imag, r[b], order=3, mode="grid-wrap"))
        assert np.abs(out[b] - ref).max() < 1e-9 * np.abs(ref).max() + 1e-12


def test_S_magnified_equals_map_coordinates():
    rng = np.random.default_rng(1)
    n, npsi = 24, 40
    sh = osh.Shift(n, npsi, n, npsi)
    c = crandn(rng, (2, npsi, npsi))
    r = np.array([[0.7, -1.9], [2.2, 0.35]])
    m = np.array([1.17, 1.43])
    out = sh.S(c, r, m)
    x, y = osh.coords(n, n, npsi, npsi, r, m)
    for b in range(2):
        yy, xx = np.meshgrid(y[b], x[b], indexing="ij")
        ref = (ndi.map_coordinates(36 * c[b].real, [yy, xx], order=3, prefilter=False, mode="grid-constant")
               + 1j * ndi.map_coordinates(36 * c[b].imag, [yy, xx], order=3, prefilter=False,
                                          mode="grid-constant"))
        assert np.abs(out[b] - ref).max() < 1e-12 * np.abs(ref).max()


def test_integer_shift_exact():
    rng = np.random.default_rng(2)
    n = 32
    sh = osh.Shift(n, n, n, n)
    psi = crandn(rng, (1, n, n))
    out = sh.curlyS(psi, np.array([[3.0, -5.0]]), np.ones(1))
    # out[ty, tx] = psi[ty - 3, tx + 5] wherever all taps are in range (periodic coeff => exact)
    inner = out[0, 3 + 2:n - 2, 2:n - 5 - 2]
    ref = psi[0, 2:n - 3 - 2, 5 + 2:n - 2]
    assert np.abs(inner - ref).max() < 1e-12


def test_coeff_inverts_circulant_and_is_self_adjoint():
    rng = np.random.default_rng(3)
    nz, n = 20, 34
    psi = crandn(rng, (nz, n))
    c = osh.coeff(psi)
    back = (4 * 4 * c + 4 * (np.roll(c, 1, 0) + np.roll(c, -1, 0) + np.roll(c, 1, 1) + np.roll(c, -1, 1))
            + (np.roll(np.roll(c, 1, 0), 1, 1) + np.roll(np.roll(c, 1, 0), -1, 1)
               + np.roll(np.roll(c, -1, 0), 1, 1) + np.roll(np.roll(c, -1, 0), -1, 1)))
    assert np.abs(back - psi).max() < 1e-12
    g = crandn(rng, (nz, n))
    assert abs(np.vdot(osh.coeff(psi), g) - np.vdot(psi, osh.coeff(g))) < 1e-12 * np.abs(np.vdot(psi, g)) + 1e-12


@pytest.mark.parametrize("real", [False, True])
def test_adjoint_S(real):
    rng = np.random.default_rng(4)
    n, npsi, nz, nzpsi = 20, 26, 18, 22
    sh = osh.Shift(n, npsi, nz, nzpsi, "float32" if real else "complex64")
    c = rng.standard_normal((3, nzpsi, npsi)) if real else crandn(rng, (3, nzpsi, npsi))
    g = rng.standard_normal((3, nz, n)) if real else crandn(rng, (3, nz, n))
    r = rng.uniform(-4, 4, (3, 2))
    m = rng.uniform(0.9, 1.4, 3)
    lhs = np.vdot(sh.S(c, r, m), g)
    rhs = np.vdot(c, sh.Sadj(g, r, m))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def _fd_setup(seed):
    rng = np.random.default_rng(seed)
    n, npsi = 20, 28
    sh = osh.Shift(n, npsi, n, npsi)
    c = crandn(rng, (2, npsi, npsi))
    r = rng.uniform(-3, 3, (2, 2))
    m = np.array([1.1, 1.25])
    return rng, sh, c, r, m


def test_dS_matches_finite_differences():
    rng, sh, c, r, m = _fd_setup(5)
    c1 = crandn(rng, c.shape)
    dr = rng.uniform(-1, 1, (2, 2))
    d = sh.dcurlySc(c, r, m, c1, dr)
    errs = []
    for h in [1e-2, 5e-3, 2.5e-3]:
        fd = (sh.S(c + h * c1, r + h * dr, m) - sh.S(c - h * c1, r - h * dr, m)) / (2 * h)
        errs.append(np.linalg.norm(fd - d))
    # central differences: error ~ h^2
    assert errs[0] / errs[1] == pytest.approx(4, rel=0.15)
    assert errs[1] / errs[2] == pytest.approx(4, rel=0.15)
    assert errs[-1] < 1e-4 * np.linalg.norm(d)


def test_d2S_is_second_differential():
    """The reference's d2s (cuda_kernels.py:437-450) subtracts w1*c1 + w2*c2,
    where w1 carries Deltar1: it pairs Deltar1 with c1.  That is the mixed second
    differential of S in the directions (c2, Deltar1) and (c1, Deltar2), which
    equals the one in (c1, Deltar1), (c2, Deltar2) only when both are equal
    (the y = z Hessian calls).  Pinned here by finite differences."""
    rng, sh, c, r, m = _fd_setup(6)
    c1, c2 = crandn(rng, c.shape), crandn(rng, c.shape)
    d1, d2 = rng.uniform(-1, 1, (2, 2)), rng.uniform(-1, 1, (2, 2))
    v = sh.d2curlySc(c, r, m, c1, d1, c2, d2)
    h = 1e-5
    fd = (sh.dcurlySc(c + h * c1, r + h * d2, m, c2, d1) - sh.dcurlySc(c - h * c1, r - h * d2, m, c2, d1)) / (2 * h)
    assert np.linalg.norm(fd - v) < 1e-6 * np.linalg.norm(v)
    # equal directions: the true second differential
    vv = sh.d2curlySc(c, r, m, c1, d1, c1, d1)
    fd = (sh.dcurlySc(c + h * c1, r + h * d1, m, c1, d1) - sh.dcurlySc(c - h * c1, r - h * d1, m, c1, d1)) / (2 * h)
    assert np.linalg.norm(fd - vv) < 1e-6 * np.linalg.norm(vv)
    # symmetric in its two directions
    w = sh.d2curlySc(c, r, m, c2, d2, c1, d1)
    assert np.linalg.norm(w - v) < 1e-12 * np.linalg.norm(v)


def test_dSadj_is_adjoint_of_dS():
    rng, sh, c, r, m = _fd_setup(7)
    c1 = crandn(rng, c.shape)
    dr = rng.uniform(-1, 1, (2, 2))
    g = crandn(rng, (2, 20, 20))
    lhs = np.sum((np.conj(sh.dcurlySc(c, r, m, c1, dr)) * g).real)
    out1, out2 = sh.dcurlySadjc(c, r, m, g)
    rhs = np.sum((np.conj(c1) * out1).real) + np.sum(dr * out2)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
