"""Pins of the Fresnel oracle (oracle/propagation.py) -- CPU only."""
import numpy as np
import pytest

from oracle import propagation as op


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


WL, VS = 1.3, 0.7
DIST = [400.0, 900.0]


@pytest.fixture(scope="module")
def prop():
    return op.Propagation(24, 16, 3, 2, WL, VS, DIST)


def test_pad_fwd_is_numpy_symmetric_and_pad_adj_its_transpose():
    rng = np.random.default_rng(0)
    f = crandn(rng, (2, 6, 10))
    assert np.array_equal(op.pad_fwd(f), np.pad(f, ((0, 0), (3, 3), (5, 5)), mode="symmetric"))
    g = crandn(rng, (2, 12, 20))
    assert abs(np.vdot(op.pad_fwd(f), g) - np.vdot(f, op.pad_adj(g))) < 1e-12 * np.abs(np.vdot(f, op.pad_adj(g)))


def test_dct_eigenfunctions(prop):
    """cos(pi k (x+1/2)/n) (x) cos(pi l (y+1/2)/nz) is an eigenvector with eigenvalue
    exp(-i pi lambda z ((k/(2n vs))^2 + (l/(2nz vs))^2)) -- fixes sign, grid and scale."""
    n, nz = 24, 16
    x = np.arange(n) + 0.5
    y = np.arange(nz) + 0.5
    for k, l in [(0, 0), (1, 0), (5, 3), (23, 15), (12, 8)]:
        v = np.cos(np.pi * l * y / nz)[:, None] * np.cos(np.pi * k * x / n)[None, :]
        for j, z in enumerate(DIST):
            lam = np.exp(-1j * np.pi * WL * z * ((k / (2 * n * VS)) ** 2 + (l / (2 * nz * VS)) ** 2))
            assert np.abs(prop.D(v, j) - lam * v).max() < 1e-12
            assert np.abs(prop.DT(v, j) - np.conj(lam) * v).max() < 1e-12


def test_constant_is_invariant(prop):
    v = np.full((16, 24), 0.3 - 0.7j)
    assert np.abs(prop.D(v, 1) - v).max() < 1e-13


def test_separable_and_dct_domain(prop):
    rng = np.random.default_rng(1)
    x = crandn(rng, (3, 16, 24))
    for j, z in enumerate(DIST):
        assert np.abs(prop.D(x, j) - op.dct_domain_D(x, WL, VS, z)).max() < 1e-12
        assert np.abs(prop.DT(x, j) - op.dct_domain_D(x, WL, VS, z, adjoint=True)).max() < 1e-12


def test_adjoint_and_conjugate_kernel(prop):
    rng = np.random.default_rng(2)
    a = crandn(rng, (3, 16, 24))
    b = crandn(rng, (3, 16, 24))
    for j in range(2):
        lhs = np.vdot(b, prop.D(a, j))
        rhs = np.vdot(prop.DT(b, j), a)
        assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # DT == D with conj(fker): compare against a propagator at negated distance
    neg = op.Propagation(24, 16, 3, 2, WL, VS, [-d for d in DIST])
    assert np.abs(prop.DT(a, 0) - neg.D(a, 0)).max() < 1e-12


def test_energy_ratio_near_one(prop):
    """||D psi||^2/||psi||^2 ~ 1 (reference test_propagation.py:59-68); exactly 1
    here because D is unitary (orthonormal DCT times a unimodular diagonal)."""
    rng = np.random.default_rng(3)
    x = crandn(rng, (2, 16, 24))
    assert np.linalg.norm(prop.D(x, 0)) / np.linalg.norm(x) == pytest.approx(1.0, abs=1e-12)


def test_ref32_floor_on_white_noise():
    """The reference builds f^2 and the chirp in float32 (propagation.py:20-28); at
    Fresnel number 1e-3 its operator sits ~1e-5 from FP64 on white noise."""
    n = 256
    fp64 = op.Propagation(n, n, 1, 1, 1.0, 1.0, [1e3])
    ref32 = op.Propagation(n, n, 1, 1, 1.0, 1.0, [1e3], ref32=True)
    rng = np.random.default_rng(4)
    x = crandn(rng, (1, n, n))
    e = np.linalg.norm(ref32.D(x, 0) - fp64.D(x, 0)) / np.linalg.norm(fp64.D(x, 0))
    assert 2e-6 < e < 1e-4
