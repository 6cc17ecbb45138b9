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


def test_usfft_constants_of_the_reference():
    """eps, mu and m as tomo.py:14-16 states them."""
    for n in (16, 256, 2048):
        eps, mu, m = ot.usfft_params(n)
        assert eps == 1e-3
        assert mu == pytest.approx(np.log(1e3) / (2 * n * n), rel=1e-15)
        assert m == 4


def _nudft_error(n, ntheta, seed=0):
    """Relative error of the gridded spectrum (R before its line transform)
    against a brute-force FP64 type-2 NUDFT of the deapodised slice at the
    polar points x0 = (t - n/2) cos / n, y0 = -(t - n/2) sin / n:
    4n sum_ij f_ij s e^{-2 pi i (x0 (j - n/2) + y0 (i - n/2))},
    s = (1 - n % 4) / (n sqrt(n ntheta)) (phi without its Gaussian factor)."""
    import scipy.fft as sfft
    theta = np.arange(ntheta) * np.pi / ntheta
    T = ot.Tomo(n, 1, theta, 0)
    rng = np.random.default_rng(seed)
    f = crandn(rng, (n, n))
    fde = np.zeros((1, 2 * n, 2 * n), complex)
    fde[:, n // 2:3 * n // 2, n // 2:3 * n // 2] = f * T.phi
    G = T.c2d * sfft.fft2(T.c2d * fde, axes=(-2, -1))
    S = T._gather(G)[:, 0, :]
    i, j = np.mgrid[0:n, 0:n]
    t = np.arange(n)
    x0 = (t[None, :] - n // 2) / n * np.cos(theta)[:, None]
    y0 = -(t[None, :] - n // 2) / n * np.sin(theta)[:, None]
    E = np.exp(-2j * np.pi * (x0[..., None, None] * (j - n / 2) + y0[..., None, None] * (i - n / 2)))
    exact = 4 * n * np.einsum("tkij,ij->tk", E, f * (1 - n % 4) / (n * np.sqrt(n * ntheta)))
    return np.linalg.norm(S - exact) / np.linalg.norm(exact)


@pytest.mark.parametrize("n", [16, 32])
def test_gridding_matches_brute_force_nudft(n, monkeypatch):
    """The stated eps / mu / m give a gridding error of 4.1-4.3e-4 against the
    exact NUDFT (inside eps = 1e-3).  The band is narrow enough that mu off by
    10% (9e-4 or 2e-4) or a half-width of 3 taps (8e-4) falls outside it."""
    err = _nudft_error(n, 12)
    assert 3.5e-4 < err < 5.0e-4, err
    orig = ot.usfft_params
    for fac in (0.9, 1.1):
        monkeypatch.setattr(ot, "usfft_params", lambda nn, f=fac: (orig(nn)[0], orig(nn)[1] * f, orig(nn)[2]))
        bad = _nudft_error(n, 12)
        assert not (3.5e-4 < bad < 5.0e-4), (fac, bad)
    monkeypatch.setattr(ot, "usfft_params", lambda nn: (orig(nn)[0], orig(nn)[1], 3))
    assert not (3.5e-4 < _nudft_error(n, 12) < 5.0e-4)


@pytest.mark.parametrize("real", [True, False])
def test_filters_have_closed_form_eigenvalues(real):
    """cos(2 pi k t / n) is an eigenvector of each detector filter with
    eigenvalue 4|k| (ramp), 4|k| sin(pi k/n)/(pi k/n) (Shepp-Logan) and
    4|k| 2 (1 - 8|k|)^3 (the reference's Parzen, u = 8|k| > 1/2 for k >= 1);
    a constant is removed by all three."""
    n, nth = 64, 3
    T = ot.Tomo(n, 2, np.arange(nth) * np.pi / nth, 0)
    t = np.arange(n)
    for k in (1, 5, 16, 31):
        v = np.cos(2 * np.pi * k * t / n) * (1.0 if real else (0.3 - 0.7j))
        data = np.broadcast_to(v, (nth, 2, n)).copy()
        for name, lam in (("ramp", 4 * k), ("shepp", 4 * k * np.sin(np.pi * k / n) / (np.pi * k / n)),
                          ("parzen", 4 * k * 2 * (1 - 8 * k) ** 3)):
            out = T.filter_sino(data, name)
            assert np.isrealobj(out) == real
            # FFT round-off in the other bins is amplified by the largest |h|
            tol = (1e-10 * abs(lam) + 1e-14 * np.abs(T.filter_h(name)).max()) * np.abs(data).max()
            assert np.abs(out - lam * data).max() < tol, (name, k)
    for name in ("ramp", "shepp", "parzen"):
        assert np.abs(T.filter_sino(np.ones((nth, 2, n)), name)).max() < 1e-9
    with pytest.raises(ValueError):
        T.filter_sino(np.ones((nth, 1, n)), "hann")


@pytest.mark.parametrize("name", ["ramp", "shepp"])
@pytest.mark.parametrize("nth", [100, 256])
def test_fbp_of_projections_is_the_scaled_object(name, nth):
    """Filtered back-projection of R of a smooth object gives the object times
    (4 / pi) sqrt(n / ntheta), plus a constant: with R = (line integrals) /
    sqrt(n ntheta), RT its adjoint, the filter 4n|f| (f in cycles per pixel)
    and the factor sqrt(n / ntheta), the continuous inversion formula
    f = int_0^pi (|nu| filtered projection) dtheta, discretised with
    dtheta = pi / ntheta, fixes that scale (the reference's FBP is the
    identity at ntheta = 16 n / pi^2); the constant is the periodic ramp
    filter's missing DC.  A dropped or doubled factor moves the scale by 2x."""
    n = 128
    T = ot.Tomo(n, 1, np.arange(nth) * np.pi / nth, 0)
    f = blob(n, n / 2 + 5, n / 2 - 3, 4.0) + 0.5 * blob(n, n / 2 - 8, n / 2 + 6, 3.0)
    rec = T.fbp(T.R(f[None]), name)[0]
    A = np.stack([f.ravel(), np.ones(f.size)], 1)
    (a, b), *_ = np.linalg.lstsq(A, rec.ravel(), rcond=None)
    assert a / (4 / np.pi * np.sqrt(n / nth)) == pytest.approx(1.0, abs=1e-2)
    assert np.linalg.norm(rec - a * f - b) < 3e-2 * np.linalg.norm(a * f)


def test_rec_tomo_is_conjugate_gradients():
    """rec_tomo's first step minimises |R (a eta) - d|^2 exactly over a (its
    derivative in a vanishes), later search directions are R-conjugate
    (Re<R eta_k, R eta_(k-1)> = 0) and the residual falls every iteration."""
    n, nth = 32, 48
    T = ot.Tomo(n, 2, np.arange(nth) * np.pi / nth, 0.9)
    rng = np.random.default_rng(5)
    d = T.R(crandn(rng, (2, n, n)) * 0.1) + 0.01 * crandn(rng, (nth, 2, n))
    trace = []
    T.rec_tomo(d, niter=5, trace=trace)
    u1, g0, e0, Re0, a0 = trace[0]
    dphi = 2 * float(np.sum(np.real(np.conj(T.R(a0 * e0) - d) * Re0)))
    assert abs(dphi) < 1e-10 * float(np.sum(np.abs(d) ** 2))
    for k in range(1, 5):
        Rk, Rk1 = trace[k][3], trace[k - 1][3]
        c = float(np.sum(np.real(np.conj(Rk) * Rk1)))
        assert abs(c) < 1e-8 * np.linalg.norm(Rk) * np.linalg.norm(Rk1)
    res = [np.linalg.norm(T.R(tr[0]) - d) for tr in trace]
    assert all(b < a for a, b in zip(res, res[1:]))
