"""Pins of the cascade oracle (oracle/cascade.py): gradient and bilinear form
against finite differences, Taylor ratios, symmetry -- CPU only."""
import numpy as np
import pytest

from oracle import cascade as oc
from oracle import propagation as op
from oracle import shift as osh


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def problem(seed=0, real=False, B=3, n=16, nobj=20, ndist=2):
    rng = np.random.default_rng(seed)
    prop = op.Propagation(n, n, B, ndist, 1.0, 1.0, [9.0, 14.0][:ndist])
    sh = osh.Shift(n, nobj, n, nobj, "float32" if real else "complex64")
    mag = np.tile(np.array([1.0, 1.12][:ndist]), (B, 1))
    cas = oc.Cascade(prop, sh, ndist, n, n, nobj, nobj, data_size=B * ndist * n * n, mag=mag, obj_real=real)
    yy, xx = np.mgrid[0:n, 0:n] - n / 2
    prb = np.stack([np.exp(-(xx ** 2 + yy ** 2) / (2 * (0.45 * n) ** 2)) * np.exp(0.3j * np.cos(0.2 * xx + j))
                    for j in range(ndist)]) * (1 + 0.05 * crandn(rng, (ndist, n, n)))
    proj = 0.4 * (rng.standard_normal((B, nobj, nobj)) if real else crandn(rng, (B, nobj, nobj)))
    pos = rng.uniform(-2, 2, (B, ndist, 2))
    d = np.abs(cas.forward(prb, proj, pos)) * (1 + 0.05 * rng.standard_normal((B, ndist, n, n)))
    # evaluate away from the data-generating point
    x = [prb * (1 + 0.1 * crandn(rng, prb.shape)), proj + 0.1 * crandn(rng, proj.shape).real * (1 if real else 1)
         + (0 if real else 0.1j * rng.standard_normal(proj.shape)), pos + rng.uniform(-0.5, 0.5, pos.shape)]
    v = [0.1 * crandn(rng, prb.shape),
         0.1 * (rng.standard_normal(proj.shape) if real else crandn(rng, proj.shape)),
         0.3 * rng.standard_normal(pos.shape)]
    return cas, x, v, d


def shifted(x, v, h):
    return [x[0] + h * v[0], x[1] + h * v[1], x[2] + h * v[2]]


def dot(g, v):
    return sum(float(np.sum(np.real(np.conj(a) * b))) for a, b in zip(g, v))


@pytest.mark.parametrize("real", [False, True])
def test_gradient_matches_finite_differences(real):
    cas, x, v, d = problem(1, real)
    g = cas.gradients(*x, d)
    f = lambda xx: cas.objective(*xx, d)  # noqa: E731
    gv = dot(g, v)
    errs = []
    for h in [1e-2, 5e-3, 2.5e-3]:
        fd = (f(shifted(x, v, h)) - f(shifted(x, v, -h))) / (2 * h)
        errs.append(abs(fd - gv))
    assert errs[-1] < 1e-5 * abs(gv)
    assert errs[0] / errs[1] == pytest.approx(4, rel=0.2)
    # each variable separately (catches compensating errors)
    for k in range(3):
        vk = [np.zeros_like(vv) for vv in v]
        vk[k] = v[k]
        h = 1e-4
        fd = (f(shifted(x, vk, h)) - f(shifted(x, vk, -h))) / (2 * h)
        assert fd == pytest.approx(dot(g, vk), rel=1e-5, abs=1e-12)


@pytest.mark.parametrize("real", [False, True])
def test_hessian_equal_directions_matches_second_differences(real):
    cas, x, v, d = problem(2, real)
    f = lambda xx: cas.objective(*xx, d)  # noqa: E731
    hvv = cas.hessian(x, v, v, d)
    h = 1e-3
    fd = (f(shifted(x, v, h)) - 2 * f(x) + f(shifted(x, v, -h))) / h ** 2
    assert fd == pytest.approx(hvv, rel=1e-4)


def test_hessian_symmetric():
    cas, x, v, d = problem(3)
    rng = np.random.default_rng(33)
    w = [0.1 * crandn(rng, v[0].shape), 0.1 * crandn(rng, v[1].shape), 0.3 * rng.standard_normal(v[2].shape)]
    assert cas.hessian(x, v, w, d) == pytest.approx(cas.hessian(x, w, v, d), rel=1e-12)


def test_taylor_ratios_F3():
    """||F3(x + h y) - F3(x) - h dF3(x)(y)|| shrinks by ~4 per halving of h
    (tests/holotomo3d/test_approximation.ipynb), and the second-order
    remainder by ~8."""
    cas, x, v, d = problem(4)
    x22 = cas.F3(x)[1]
    _, dy = cas.dF3(x, v)
    d2 = cas.d2F_dF3(x, v, v, [None, None, None])[1]
    e1, e2 = [], []
    for h in [0.1, 0.05, 0.025]:
        xh = cas.F3(shifted(x, v, h))[1]
        e1.append(np.linalg.norm(xh - x22 - h * dy[1]))
        e2.append(np.linalg.norm(xh - x22 - h * dy[1] - 0.5 * h * h * d2))
    assert e1[0] / e1[1] == pytest.approx(4, rel=0.15)
    assert e1[1] / e1[2] == pytest.approx(4, rel=0.15)
    assert e2[0] / e2[1] == pytest.approx(8, rel=0.2)


def test_single_precision_mode_tracks_fp64():
    """The oracle's single-precision mode (used only to time the reference's
    precision on the CPU) computes in complex64 and stays within FP32 rounding
    of the FP64 oracle: gradient relative L2 <= 1e-4."""
    from oracle import cascade as oc
    from oracle import propagation as op
    from oracle import shift as osh
    n, nobj, nd, B = 32, 24, 2, 2
    rng = np.random.default_rng(77)
    mag = rng.uniform(1.0, 1.1, (B, nd))
    dist = [40.0, 55.0]
    out = []
    for single in (False, True):
        cas = oc.Cascade(op.Propagation(n, n, B, nd, 1.0, 1.0, dist, dtype=np.complex64 if single else np.complex128),
                         osh.Shift(n, nobj, n, nobj, single=single), nd, n, n, nobj, nobj, B * nd * n * n, mag,
                         rho=(1.0, 0.5, 0.2), single=single)
        rng = np.random.default_rng(78)
        prb = np.exp(0.2j * rng.standard_normal((nd, n, n)))
        proj = 0.3 * (rng.standard_normal((B, nobj, nobj)) + 1j * rng.standard_normal((B, nobj, nobj)))
        pos = rng.uniform(-2, 2, (B, nd, 2))
        d = rng.uniform(0.5, 1.5, (B, nd, n, n))
        cast = (lambda a: a.astype(np.complex64)) if single else (lambda a: a)
        out.append(cas.gradients(cast(prb), cast(proj), pos, d.astype(np.float32) if single else d))
    for g64, g32 in zip(*out):
        assert np.linalg.norm(g32 - g64) / np.linalg.norm(g64) < 1e-4
    assert out[1][0].dtype == np.complex64 and out[1][1].dtype == np.complex64
