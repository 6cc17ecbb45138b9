"""Pins of the back-shift oracle (oracle/shift.py: sback, sback_forward,
coeffback) -- CPU only.

* a literal loop over detector pixels and window offsets, in the order of
  the sback kernel (cuda_kernels.py:327-356), equals the sparse form;
* m = 1: the window is j in [ceil(d - 2), (int)(d + 2)), i.e. the plain
  spline's taps -1..1 without the last one (weight d^3), so Sback equals the
  pinned Sadj minus that tap, and coeffback's filter is the [1 4 1]
  prefilter, so coeffback equals coeff;
* Sback is the transpose of the sback gather for any m;
* for any m with D > 0, coeffback inverts the periodic convolution with the
  sampled stretched spline taps phi(j/m), |j| <= 4 (closed form)."""
import numpy as np
import pytest

from oracle import shift as osh


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def sback_loops(g, r, m, n, npsi, nz, nzpsi):
    """Literal restatement of the sback scatter loop (dir 1)."""
    out = np.zeros((g.shape[0], nzpsi, npsi), dtype=np.complex128)
    for b in range(g.shape[0]):
        half = (m[b] - 1) / 2
        span = 2 * m[b]
        for ty in range(nz):
            for tx in range(n):
                x = m[b] * (tx - n // 2) - r[b, 1] + half + npsi // 2
                y = m[b] * (ty - nz // 2) - r[b, 0] + half + nzpsi // 2
                ix, iy = int(np.floor(x)), int(np.floor(y))
                dx, dy = x - ix, y - iy
                for jy in range(int(np.ceil(dy - span)), int(dy + span)):
                    if not 0 <= iy + jy < nzpsi:
                        continue
                    for jx in range(int(np.ceil(dx - span)), int(dx + span)):
                        if 0 <= ix + jx < npsi:
                            out[b, iy + jy, ix + jx] += (osh.phi((dx - jx) / m[b]) * osh.phi((dy - jy) / m[b])
                                                         * g[b, ty, tx])
    return out


def test_literal_loops_and_m1_reduction():
    rng = np.random.default_rng(0)
    n, npsi, nz, nzpsi, B = 12, 16, 10, 14, 2
    r = rng.uniform(-2, 2, (B, 2))
    g = crandn(rng, (B, nz, n))
    for m in (np.array([1.0, 1.0]), np.array([1.4, 0.8])):
        assert np.allclose(osh.sback(g, r, m, n, npsi, nz, nzpsi), sback_loops(g, r, m, n, npsi, nz, nzpsi),
                           atol=1e-12)
    # m = 1: Sadj minus the j = 2 tap of each axis
    o = osh.Shift(n, npsi, nz, nzpsi)
    m = np.ones(B)
    x, y = osh.coords(n, nz, npsi, nzpsi, r, m)

    def drop_last(npix, npsi_, coord):
        W = osh.axis_matrix(npix, npsi_, coord).tolil()
        i0 = np.floor(coord).astype(int)
        for t in range(npix):
            if 0 <= i0[t] + 2 < npsi_:
                W[t, i0[t] + 2] = 0.0
        return W.tocsr()

    ref = np.stack([drop_last(nz, nzpsi, y[b]).T @ g[b] @ drop_last(n, npsi, x[b]) for b in range(B)])
    assert np.allclose(osh.sback(g, r, m, n, npsi, nz, nzpsi), ref, atol=1e-12)
    full = o.Sadj(g, r, m)
    assert np.linalg.norm(full - ref) > 1e-3 * np.linalg.norm(full)  # the dropped tap matters
    psi = crandn(rng, (B, nzpsi, npsi))
    assert np.allclose(osh.coeffback(psi, 1.0), o.coeff(psi), atol=1e-12)


@pytest.mark.parametrize("m", [0.7, 1.3, 2.2])
def test_sback_is_transpose_of_gather(m):
    rng = np.random.default_rng(1)
    n, npsi, nz, nzpsi, B = 16, 40, 18, 36, 2
    r = rng.uniform(-2, 2, (B, 2))
    g = crandn(rng, (B, nz, n))
    c = crandn(rng, (B, nzpsi, npsi))
    lhs = np.vdot(osh.sback_forward(c, r, m, n, npsi, nz, nzpsi), g)
    rhs = np.vdot(c, osh.sback(g, r, m, n, npsi, nz, nzpsi))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


@pytest.mark.parametrize("m", [1.3, 1.6])
def test_coeffback_inverts_sampled_stretched_kernel(m):
    rng = np.random.default_rng(2)
    N = 32
    d = osh.coeffback_filter(N, m)
    assert d.min() > 0
    c = crandn(rng, (1, N, N))
    taps = {j: osh.phi(j / m) for j in range(-4, 5)}
    conv = sum(taps[j] * np.roll(c, j, axis=-1) for j in taps)
    conv = sum(taps[j] * np.roll(conv, j, axis=-2) for j in taps)
    assert np.allclose(osh.coeffback(conv, m), c, atol=1e-10)
