"""Pins of the BH-loop oracle (oracle/bh.py) -- CPU only.

The Laplacian against scipy.ndimage, the biharmonic as the gradient of the
padded energy, the first step as the Newton step of the objective along the
direction (central differences), Hessian conjugacy of successive directions,
and descent of the objective."""
import numpy as np
import pytest
import scipy.ndimage as nd

from oracle import bh as ob


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def lap_scipy(u):
    f = lambda a: nd.laplace(a, mode=["constant", "wrap", "wrap"])  # noqa: E731
    return f(u.real) + 1j * f(u.imag) if np.iscomplexobj(u) else f(u)


def test_lap_matches_scipy():
    rng = np.random.default_rng(0)
    u = crandn(rng, (5, 6, 7))
    p = ob.pad2(u)
    assert np.abs(ob.lap(p[1:-3], p[2:-2], p[3:-1]) - lap_scipy(u)).max() < 1e-12
    assert np.abs(ob.biharmonic(u) - lap_scipy(lap_scipy(ob.pad2(u)))[2:-2]).max() < 1e-11
    # interior slices: plain lap(lap(u)) with zero outside the volume
    assert np.abs(ob.biharmonic(u)[1:-1] - lap_scipy(lap_scipy(u))[1:-1]).max() < 1e-11


def test_biharmonic_is_gradient_of_padded_energy():
    rng = np.random.default_rng(1)
    u, v = crandn(rng, (4, 5, 6)), crandn(rng, (4, 5, 6))

    def energy(w):
        p = ob.pad2(w)
        L = ob.lap(p[:-2], p[1:-1], p[2:])  # slices -1 .. nz of the padded volume
        return float(np.sum(np.abs(L) ** 2))

    h = 1e-4
    fd = (energy(u + h * v) - energy(u - h * v)) / (2 * h)
    g = 2 * ob.biharmonic(u)
    assert fd == pytest.approx(float(np.sum(np.real(np.conj(g) * v))), rel=1e-7)
    # hessian_laplacian is symmetric (the biharmonic is self-adjoint)
    s = ob.BHSolver.__new__(ob.BHSolver)
    s.lam_laplacian, s.obj_size = 0.7, 120
    assert s.hessian_laplacian(u, v) == pytest.approx(s.hessian_laplacian(v, u), rel=1e-12)


def problem(seed=0, lam_prbfit=0.0, lam_laplacian=0.0, n=16, nobj=8, ntheta=6, ndist=2):
    rng = np.random.default_rng(seed)
    theta = np.arange(ntheta) * np.pi / ntheta
    mag = np.tile(np.array([1.0, 0.9][:ndist]), (ntheta, 1))
    s = ob.BHSolver(n, n, nobj, nobj, ndist, 1.0, 1.0, [9.0, 14.0][:ndist], theta, mag,
                    rho=(1.0, 0.5, 0.3), lam_prbfit=lam_prbfit, lam_laplacian=lam_laplacian)
    yy, xx = np.mgrid[0:n, 0:n] - n / 2
    prb = np.stack([np.exp(-(xx ** 2 + yy ** 2) / (2 * (0.45 * n) ** 2)) * np.exp(0.3j * np.cos(0.2 * xx + j))
                    for j in range(ndist)])
    zz, yo, xo = np.mgrid[0:nobj, 0:nobj, 0:nobj] - nobj / 2
    obj = 0.3 * np.exp(-(zz ** 2 + yo ** 2 + xo ** 2) / (2 * (0.25 * nobj) ** 2)) * (1 + 0.5j)
    pos = rng.uniform(-1, 1, (ntheta, ndist, 2))
    truth = {"obj": obj, "prb": prb, "pos": pos}
    s.data = s.gen_sqrt_data(truth)
    s.ref = s.gen_sqrt_ref(prb) * 1.02
    start = {"obj": obj * 0.7 + 0.02 * crandn(rng, obj.shape), "prb": prb * (1 + 0.05 * crandn(rng, prb.shape)),
             "pos": pos + 0.3 * rng.standard_normal(pos.shape)}
    return s, start


def test_first_step_is_newton_step_of_objective():
    s, start = problem(1, lam_prbfit=0.3)
    trace = []
    s.BH(start, 1, trace=trace)
    t = trace[0]
    x, e = t["vars"], t["etas"]

    def phi(h):
        return s.min({k: x[k] + h * e[k] for k in x})

    h = 1e-3 * abs(t["alpha"])  # the step's own scale: small truncation and cancellation error
    p0, pp, pm = phi(0.0), phi(h), phi(-h)
    d1 = (pp - pm) / (2 * h)
    d2 = (pp - 2 * p0 + pm) / h ** 2
    assert t["beta"] == 0
    assert t["alpha"] == pytest.approx(-d1 / d2, rel=1e-6)
    # the direction is the negative preconditioned gradient, so phi'(0) < 0
    assert d1 < 0


def test_directions_are_hessian_conjugate():
    s, start = problem(2, lam_prbfit=0.2, lam_laplacian=0.05)
    trace = []
    s.BH(start, 2, trace=trace)
    x1, e0, e1, g1 = trace[1]["vars"], trace[0]["etas"], trace[1]["etas"], trace[1]["grads"]
    assert trace[1]["beta"] != 0
    # H(e1, e0) = beta H(e0, e0) - H(g, e0) vanishes up to the reference's
    # d2s pairing (its second-order shift term is exactly bilinear only for
    # equal directions), which leaves about 1e-6 of the terms it cancels
    assert abs(s.hessian(x1, e1, e0)) < 1e-5 * abs(s.hessian(x1, g1, e0))


def test_objective_decreases():
    s, start = problem(3, lam_prbfit=0.1, lam_laplacian=0.02)
    x0 = dict(start)
    x0n = {k: np.array(v) for k, v in x0.items()}
    x0n["obj"] = x0n["obj"] / s.norm_const
    x0n["proj"] = s.fwd_tomo(x0n["obj"])
    f0 = s.min(x0n)
    _, hist = s.BH(start, 4)
    errs = [f0] + [h[2] for h in hist]
    assert all(b < a for a, b in zip(errs, errs[1:]))
    assert errs[-1] < 0.5 * errs[0]
