"""GPU parity of the USFFT Radon transform R / RT against the FP64 oracle.

Tolerance: relative L2 <= 1e-5 (FP32 kernels vs FP64 oracle; the angles are
rounded to float32 on both sides, as the reference stores them).  At the
bench size (256^3 slab share, 384 angles) the checks are size-independent:
adjointness of RT to R, linearity, and the projection trajectory of an
off-centre blob.
"""
import numpy as np
import pytest
import torch

from oracle import tomo as ot

pytestmark = pytest.mark.gpu
REL_TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b))


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def data(rng, shape, real):
    x = rng.standard_normal(shape)
    if not real:
        x = x + 1j * rng.standard_normal(shape)
    return x.astype(np.float32 if real else np.complex64)


def angles(ntheta, span=np.pi):
    return (np.arange(ntheta) * span / ntheta).astype(np.float32).astype(np.float64)


CASES = [  # n, nz, ntheta, mask_r, nz used
    (8, 2, 5, 0, 2),
    (16, 3, 12, 0, 3),
    (32, 2, 24, 0.9, 2),
    (64, 4, 48, 0, 3),
    (128, 2, 30, 0.8, 2),
    (32, 32, 20, 0, 32),  # whole 16-slice blocks in the adjoint's gather
]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("real", [True, False])
def test_R_matches_oracle(gpu, case, real):
    from holotomocupy_amd import Tomo
    n, nz, ntheta, mask_r, nzu = case
    print(f"[case] tomo R n={n} nz={nzu} ntheta={ntheta} real={real}", flush=True)
    rng = np.random.default_rng(n + ntheta)
    th = angles(ntheta)
    g, o = Tomo(n, nz, th, mask_r), ot.Tomo(n, nz, th, mask_r)
    x = data(rng, (nzu, n, n), real)
    y = g.R(dev(x))
    torch.cuda.synchronize()
    assert y.dtype == (torch.float32 if real else torch.complex64) and tuple(y.shape) == (ntheta, nzu, n)
    assert rel(y.cpu().numpy(), o.R(x.astype(np.float64 if real else np.complex128))) < REL_TOL


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("real", [True, False])
def test_RT_matches_oracle(gpu, case, real):
    from holotomocupy_amd import Tomo
    n, nz, ntheta, mask_r, nzu = case
    print(f"[case] tomo RT n={n} nz={nzu} ntheta={ntheta} real={real}", flush=True)
    rng = np.random.default_rng(7 * n + ntheta)
    th = angles(ntheta)
    g, o = Tomo(n, nz, th, mask_r), ot.Tomo(n, nz, th, mask_r)
    d = data(rng, (ntheta, nzu, n), real)
    x = g.RT(dev(d))
    torch.cuda.synchronize()
    assert tuple(x.shape) == (nzu, n, n)
    assert rel(x.cpu().numpy(), o.RT(d.astype(np.float64 if real else np.complex128))) < REL_TOL


def test_nonuniform_angles_and_full_circle(gpu):
    from holotomocupy_amd import Tomo
    print("[case] tomo random angles n=32", flush=True)
    rng = np.random.default_rng(3)
    th = np.sort(rng.uniform(-np.pi, 2 * np.pi, 17)).astype(np.float32).astype(np.float64)
    g, o = Tomo(32, 1, th, 0), ot.Tomo(32, 1, th, 0)
    x = data(rng, (1, 32, 32), False)
    assert rel(g.R(dev(x)).cpu().numpy(), o.R(x.astype(np.complex128))) < REL_TOL
    d = data(rng, (17, 1, 32), False)
    assert rel(g.RT(dev(d)).cpu().numpy(), o.RT(d.astype(np.complex128))) < REL_TOL


def test_bench_size_adjoint_linearity_trajectory(gpu):
    """256^3 share with 384 angles: <R x, y> = <x, RT y>, R linear, blob on
    its sinusoid (size-independent properties; the oracle is too slow here)."""
    from holotomocupy_amd import Tomo
    n, nz, ntheta = 256, 4, 384
    print(f"[case] tomo bench-size n={n} nz={nz} ntheta={ntheta}", flush=True)
    th = angles(ntheta)
    g = Tomo(n, nz, th, 0)
    gen = torch.Generator(device="cuda").manual_seed(5)
    x = torch.randn((nz, n, n), dtype=torch.complex64, device="cuda", generator=gen)
    x2 = torch.randn((nz, n, n), dtype=torch.complex64, device="cuda", generator=gen)
    y = torch.randn((ntheta, nz, n), dtype=torch.complex64, device="cuda", generator=gen)
    Rx, RTy = g.R(x), g.RT(y)
    lhs = torch.vdot(Rx.flatten().to(torch.complex128), y.flatten().to(torch.complex128))
    rhs = torch.vdot(x.flatten().to(torch.complex128), RTy.flatten().to(torch.complex128))
    assert abs(complex(lhs - rhs)) < 1e-5 * abs(complex(lhs))
    lin = g.R(2.0 * x - 0.5j * x2) - (2.0 * Rx - 0.5j * g.R(x2))
    assert float(torch.linalg.norm(lin) / torch.linalg.norm(Rx)) < 1e-5
    i, j = np.mgrid[0:n, 0:n]
    ic, jc = n / 2 + 40.0, n / 2 - 25.0
    blob = np.exp(-((i - ic) ** 2 + (j - jc) ** 2) / (2 * 3.0 ** 2)).astype(np.float32)
    p = g.R(dev(blob[None].repeat(nz, 0))).cpu().numpy()[:, 0].astype(np.float64)
    w = np.abs(p) ** 4
    t_peak = (w * np.arange(n)).sum(1) / w.sum(1)
    t_ref = n / 2 + (jc - n / 2) * np.cos(th) - (ic - n / 2) * np.sin(th)
    assert np.abs(t_peak - t_ref).max() < 0.2


def test_rejects_bad_sizes(gpu):
    from holotomocupy_amd import Tomo
    from holotomocupy_amd._lib import HtamdError
    with pytest.raises(HtamdError):
        Tomo(47, 1, angles(4), 0)  # odd widths are refused (any even width up to 4096 is taken)
    with pytest.raises(HtamdError):
        Tomo(4098, 1, angles(4), 0)
    g = Tomo(16, 2, angles(4), 0)
    with pytest.raises(ValueError):
        g.R(torch.zeros((3, 16, 16), device="cuda"))


def test_fbp_and_rec_tomo_shapes(gpu):
    from holotomocupy_amd import Tomo
    print("[case] tomo fbp / rec_tomo n=32", flush=True)
    n, ntheta = 32, 48
    th = angles(ntheta)
    g = Tomo(n, 1, th, 0)
    i, j = np.mgrid[0:n, 0:n]
    f = np.exp(-((i - n / 2 - 3) ** 2 + (j - n / 2 + 4) ** 2) / (2 * 3.0 ** 2)).astype(np.float32)[None]
    d = g.R(dev(f))
    rec = g.fbp(d, "shepp")
    assert tuple(rec.shape) == (1, n, n) and rec.dtype == torch.float32
    u = g.rec_tomo(d, 8)
    r0 = float(torch.linalg.norm(d))
    assert float(torch.linalg.norm(g.R(u) - d)) < 0.2 * r0


def test_R_RT_into_given_output(gpu):
    """R / RT written into a caller's contiguous tensor equal the allocating
    calls bit for bit; a wrong shape, dtype or a non-contiguous view is
    rejected before any launch."""
    from holotomocupy_amd import Tomo
    print("[case] R/RT into a given output", flush=True)
    n, nth, nz = 32, 12, 3
    th = np.arange(nth) * np.pi / nth
    g = Tomo(n, nz, th, 0.9)
    gen = torch.Generator(device="cuda").manual_seed(5)
    u = torch.randn((nz, n, n), dtype=torch.complex64, device="cuda", generator=gen)
    d = torch.randn((nth, nz, n), dtype=torch.complex64, device="cuda", generator=gen)
    out = torch.empty((nth, nz, n), dtype=torch.complex64, device="cuda")
    assert g.R(u, out=out) is out and torch.equal(out, g.R(u))
    back = torch.empty((nz, n, n), dtype=torch.complex64, device="cuda")
    assert g.RT(d, out=back) is back and torch.equal(back, g.RT(d))
    with pytest.raises(ValueError):
        g.R(u, out=torch.empty((nth, nz, n + 1), dtype=torch.complex64, device="cuda"))
    with pytest.raises(ValueError):
        g.R(u, out=torch.empty((nth, nz, n), dtype=torch.float32, device="cuda"))
    with pytest.raises(ValueError):
        g.RT(d, out=torch.empty((n, nz, n), dtype=torch.complex64, device="cuda").transpose(0, 1))


@pytest.mark.parametrize("name", ["ramp", "shepp", "parzen"])
@pytest.mark.parametrize("n,ntheta,nz,cplx", [(64, 96, 2, True), (128, 64, 3, False), (256, 32, 1, True)])
def test_filter_and_fbp_match_oracle(gpu, name, n, ntheta, nz, cplx):
    """_filter_sino (one LDS pass per line) and fbp (the filter folded into
    RT's line transform) against the FP64 oracle restatement of
    tomo.py:118-186, relative L2 <= 1e-5; a complex NumPy sinogram keeps its
    imaginary part."""
    from holotomocupy_amd import Tomo
    from oracle import tomo as ot
    print(f"[case] filter/fbp {name} n={n} ntheta={ntheta} nz={nz} complex={cplx}", flush=True)
    rng = np.random.default_rng(n + ntheta)
    th = angles(ntheta)
    d = rng.standard_normal((ntheta, nz, n))
    if cplx:
        d = d + 1j * rng.standard_normal((ntheta, nz, n))
    d32 = d.astype(np.complex64 if cplx else np.float32)
    g, o = Tomo(n, nz, th, 0.9), ot.Tomo(n, nz, th, 0.9)
    f = g._filter_sino(dev(d32), name)
    fr = o.filter_sino(d32.astype(np.complex128 if cplx else np.float64), name)
    assert f.dtype == (torch.complex64 if cplx else torch.float32)
    assert rel(f.cpu().numpy(), fr) < 1e-5, name
    rec = g.fbp(d32, name)  # host array in: dtype kept
    assert rec.dtype == (torch.complex64 if cplx else torch.float32) and tuple(rec.shape) == (nz, n, n)
    ref = o.fbp(d32.astype(np.complex128 if cplx else np.float64), name)
    err = rel(rec.cpu().numpy(), ref)
    print(f"[fbp] {name} n={n}: {err:.2e}", flush=True)
    assert err < 1e-5, name
    # the fused pass equals the filter followed by RT
    assert rel(rec.cpu().numpy(), (g.RT(f) * float(np.sqrt(n / ntheta))).cpu().numpy()) < 1e-5
    with pytest.raises(ValueError):
        g.fbp(d32, "hann")


def test_rec_tomo_matches_oracle(gpu):
    """CG iterations of rec_tomo (tomo.py:188-210) against the FP64 oracle's.
    After four iterations the iterate is compared at 1e-4 relative L2 (each
    iteration adds the FP32 error of one R and one RT, ~5e-6 each).  Later
    iterations are ill-conditioned: on the CPU, perturbing the oracle's R and
    RT by 5e-6 relative noise moves its 8th iterate by 4e-2 and its objective
    by 4e-3, so after eight iterations the test checks the objective
    |R u - d|^2: decreasing every iteration and within 1e-2 of the oracle's."""
    from holotomocupy_amd import Tomo
    from oracle import tomo as ot
    print("[case] rec_tomo 4 / 8 iterations n=64", flush=True)
    n, ntheta, nz = 64, 96, 2
    th = angles(ntheta)
    rng = np.random.default_rng(9)
    g, o = Tomo(n, nz, th, 0.9), ot.Tomo(n, nz, th, 0.9)
    i, j = np.mgrid[0:n, 0:n]
    f = np.exp(-((i - n / 2 - 3) ** 2 + (j - n / 2 + 4) ** 2) / (2 * 5.0 ** 2))
    obj = np.stack([f, 0.5 * f.T]) * (1 + 0.1j)
    d = o.R(obj) + 0.01 * (rng.standard_normal((ntheta, nz, n)) + 1j * rng.standard_normal((ntheta, nz, n)))
    d32 = d.astype(np.complex64)
    d64 = d32.astype(np.complex128)
    u4 = g.rec_tomo(dev(d32), 4)
    err = rel(u4.cpu().numpy(), o.rec_tomo(d64, 4))
    print(f"[rec_tomo] 4 iterations: {err:.2e}", flush=True)
    assert err < 1e-4
    objs = []
    for it in range(1, 9):
        u = g.rec_tomo(dev(d32), it).cpu().numpy().astype(np.complex128)
        objs.append(np.linalg.norm(o.R(u) - d64) ** 2)
    assert all(b < a for a, b in zip(objs, objs[1:]))
    ref8 = np.linalg.norm(o.R(o.rec_tomo(d64, 8)) - d64) ** 2
    print(f"[rec_tomo] objective after 8 iterations: {objs[-1]:.6e} (oracle {ref8:.6e})", flush=True)
    assert abs(objs[-1] - ref8) / ref8 < 1e-2
