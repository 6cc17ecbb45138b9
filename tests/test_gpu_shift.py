"""GPU parity of the B-spline shift family against the FP64 oracle.

Tolerance: relative L2 <= 1e-5 (FP32 kernels vs FP64 oracle); adjoint
dot-product tests <= 1e-5.
"""
import numpy as np
import pytest
import torch

from oracle import shift as osh

pytestmark = pytest.mark.gpu
REL_TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b))


def data(rng, shape, real):
    x = rng.standard_normal(shape)
    if not real:
        x = x + 1j * rng.standard_normal(shape)
    return x.astype(np.float32 if real else np.complex64)


GEOMS = [  # n, npsi, nz, nzpsi, B
    (32, 32, 32, 32, 3),
    (24, 40, 20, 36, 2),
    (64, 48, 48, 64, 2),
    (130, 96, 70, 100, 2),
]


def setup(geom, real, seed, mrange=(0.9, 1.4), rmax=6.0):
    from holotomocupy_amd import Shift
    n, npsi, nz, nzpsi, B = geom
    rng = np.random.default_rng(seed)
    dt = "float32" if real else "complex64"
    g = Shift(n, npsi, nz, nzpsi, dt)
    o = osh.Shift(n, npsi, nz, nzpsi, dt)
    r = rng.uniform(-rmax, rmax, (B, 2)).astype(np.float32)
    m = rng.uniform(*mrange, B).astype(np.float32)
    return rng, g, o, r, m


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


@pytest.mark.parametrize("geom", GEOMS)
@pytest.mark.parametrize("real", [False, True])
def test_S_Sadj_coeff(gpu, geom, real):
    print(f"[case] S/Sadj/coeff geom={geom} real={real}", flush=True)
    rng, g, o, r, m = setup(geom, real, 11)
    n, npsi, nz, nzpsi, B = geom
    c = data(rng, (B, nzpsi, npsi), real)
    s = data(rng, (B, nz, n), real)
    assert rel(g.S(dev(c), dev(r), dev(m)).cpu(), o.S(c, r, m)) < REL_TOL
    assert rel(g.Sadj(dev(s), dev(r), dev(m)).cpu(), o.Sadj(s, r, m)) < REL_TOL
    assert rel(g.coeff(dev(c)).cpu(), o.coeff(c)) < REL_TOL
    assert rel(g.curlyS(dev(c), dev(r), dev(m)).cpu(), o.curlyS(c, r, m)) < REL_TOL


@pytest.mark.parametrize("real", [False, True])
def test_S_Sadj_wide_magnifications(gpu, real):
    """S walks the coefficient rows of 8 output rows at once and Sadj the union
    of 4 outputs' candidate detector rows: magnifications far from 1 widen
    those walks (2.3: 8-row windows spread over ~20 coefficient rows; 0.45:
    many outputs share a row), and 70 and 20 detector rows end mid-block."""
    from holotomocupy_amd import Shift
    print(f"[case] S/Sadj wide magnifications real={real}", flush=True)
    rng = np.random.default_rng(15)
    dt = "float32" if real else "complex64"
    for n, npsi, nz, nzpsi in ((64, 160, 70, 170), (96, 48, 20, 40)):
        B = 4
        g, o = Shift(n, npsi, nz, nzpsi, dt), osh.Shift(n, npsi, nz, nzpsi, dt)
        r = rng.uniform(-4, 4, (B, 2)).astype(np.float32)
        m = np.array([0.45, 1.0, 1.7, 2.3], dtype=np.float32)
        c = data(rng, (B, nzpsi, npsi), real)
        s = data(rng, (B, nz, n), real)
        assert rel(g.S(dev(c), dev(r), dev(m)).cpu(), o.S(c, r, m)) < REL_TOL, (n, npsi, nz, nzpsi)
        assert rel(g.Sadj(dev(s), dev(r), dev(m)).cpu(), o.Sadj(s, r, m)) < REL_TOL, (n, npsi, nz, nzpsi)


@pytest.mark.parametrize("geom", GEOMS[:3])
@pytest.mark.parametrize("real", [False, True])
def test_derivatives(gpu, geom, real):
    print(f"[case] dS/d2S/dSadj geom={geom} real={real}", flush=True)
    rng, g, o, r, m = setup(geom, real, 12)
    n, npsi, nz, nzpsi, B = geom
    c, c1, c2 = (data(rng, (B, nzpsi, npsi), real) for _ in range(3))
    d1 = rng.uniform(-1, 1, (B, 2)).astype(np.float32)
    d2 = rng.uniform(-1, 1, (B, 2)).astype(np.float32)
    s = data(rng, (B, nz, n), real)
    assert rel(g.dcurlySc(dev(c), dev(r), dev(m), dev(c1), dev(d1)).cpu(),
               o.dcurlySc(c, r, m, c1, d1)) < REL_TOL
    assert rel(g.d2curlySc(dev(c), dev(r), dev(m), dev(c1), dev(d1), dev(c2), dev(d2)).cpu(),
               o.d2curlySc(c, r, m, c1, d1, c2, d2)) < REL_TOL
    a1, a2 = g.dcurlySadjc(dev(c), dev(r), dev(m), dev(s))
    b1, b2 = o.dcurlySadjc(c, r, m, s)
    assert rel(a1.cpu(), b1) < REL_TOL
    assert rel(a2.cpu(), b2) < REL_TOL


def test_strided_shift_views_and_2d(gpu):
    """r = pos[:, j] and m = mag[:, j] views are consumed with their strides."""
    print("[case] strided views", flush=True)
    rng, g, o, _, _ = setup((32, 32, 32, 32, 4), False, 13)
    c = data(rng, (4, 32, 32), False)
    pos = rng.uniform(-3, 3, (4, 3, 2)).astype(np.float32)
    mag = rng.uniform(1.0, 1.3, (4, 3)).astype(np.float32)
    pos_d, mag_d = dev(pos), dev(mag)
    for j in range(3):
        out = g.S(dev(c), pos_d[:, j], mag_d[:, j]).cpu()
        assert rel(out, o.S(c, pos[:, j], mag[:, j])) < REL_TOL
    # scalar magnification
    assert rel(g.S(dev(c), pos_d[:, 0], 1.0).cpu(), o.S(c, pos[:, 0], np.ones(4))) < REL_TOL


def test_adjoint_identities_large(gpu):
    """C2-sized shift: 256^2, 32 angles, shifts U[-5, 5], adjoint dot tests.
    Each identity <A x, y> = <x, A* y> is checked relative to ||A x|| ||y||
    (the Cauchy-Schwarz scale): for random x, y the inner product itself can
    be far smaller than that, and FP32 rounding is relative to the scale."""
    print("[case] adjoint 256^2 x 32", flush=True)
    from holotomocupy_amd import Shift
    rng = np.random.default_rng(1002)
    gen = torch.Generator(device="cuda").manual_seed(1002)
    B, n = 32, 256
    g = Shift(n, n, n, n, "complex64")
    r = dev(rng.uniform(-5, 5, (B, 2)).astype(np.float32))
    m = torch.ones(B, device="cuda")
    c = torch.randn(B, n, n, dtype=torch.complex64, device="cuda", generator=gen)
    s = torch.randn(B, n, n, dtype=torch.complex64, device="cuda", generator=gen)
    d128 = lambda t: t.flatten().to(torch.complex128)  # noqa: E731
    Sc = g.S(c, r, m)
    lhs = torch.vdot(d128(Sc), d128(s))
    rhs = torch.vdot(d128(c), d128(g.Sadj(s, r, m)))
    assert abs(lhs - rhs) / (torch.linalg.norm(d128(Sc)) * torch.linalg.norm(d128(s))) < 1e-6
    Cc = g.coeff(c)
    lhs = torch.vdot(d128(Cc), d128(s))
    rhs = torch.vdot(d128(c), d128(g.coeff(s)))
    assert abs(lhs - rhs) / (torch.linalg.norm(d128(Cc)) * torch.linalg.norm(d128(s))) < 1e-6
    # dSadj identity: <dS(c1, dr), g> = <c1, out1> + dr . out2
    c1 = torch.randn(c.shape, dtype=torch.complex64, device="cuda", generator=gen)
    dr = dev(rng.uniform(-1, 1, (B, 2)).astype(np.float32))
    cc = g.coeff(c)
    dS = g.dcurlySc(cc, r, m, c1, dr)
    lhs = (d128(dS).conj() * d128(s)).real.sum()
    o1, o2 = g.dcurlySadjc(cc, r, m, s)
    rhs = (d128(c1).conj() * d128(o1)).real.sum() + (dr.double() * o2.double()).sum()
    assert abs(lhs - rhs) / (torch.linalg.norm(d128(dS)) * torch.linalg.norm(d128(s))) < 1e-6


def test_deterministic(gpu):
    print("[case] determinism", flush=True)
    rng, g, o, r, m = setup((64, 48, 48, 64, 2), False, 14)
    s = dev(data(rng, (2, 48, 64), False))
    c = dev(data(rng, (2, 64, 48), False))
    a = g.Sadj(s, dev(r), dev(m))
    b = g.Sadj(s, dev(r), dev(m))
    assert torch.equal(a, b)
    x1 = g.dcurlySadjc(c, dev(r), dev(m), s)
    x2 = g.dcurlySadjc(c, dev(r), dev(m), s)
    assert torch.equal(x1[1], x2[1])


@pytest.mark.parametrize("shape", [(2, 300, 520), (3, 6, 256), (1, 1024, 1024), (2, 38, 300)])
@pytest.mark.parametrize("real", [False, True])
def test_coeff_wide_planes(gpu, shape, real):
    """Planes at least 256 wide take the column-strip prefilter: partial
    panels in x, partial strips in y, planes shorter than the filter reach."""
    from holotomocupy_amd import Shift
    print(f"[case] coeff wide planes shape={shape} real={real}", flush=True)
    B, nzpsi, npsi = shape
    rng = np.random.default_rng(13)
    c = data(rng, shape, real)
    g = Shift(npsi, npsi, nzpsi, nzpsi, "float32" if real else "complex64")
    assert rel(g.coeff(dev(c)).cpu(), osh.coeff(c, real=real)) < REL_TOL
