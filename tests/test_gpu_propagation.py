"""GPU parity of the Fresnel propagator against the FP64 oracle.

Tolerances: relative L2 <= 1e-5 against the FP64 oracle (the kernels compute
in FP32 with FP64-generated chirp and twiddle tables); adjoint dot-product
test <= 1e-5 (reference test: tests/propagation/test_propagation.py:38-53).
"""
import numpy as np
import pytest
import torch

from oracle import propagation as orc

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def crandn(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def make(n, nz, ntheta, fresnel):
    from holotomocupy_amd import Propagation
    # voxelsize = wavelength = 1, distance = 1/N_F  ->  phase pi (fx^2+fy^2) / N_F
    dist = [1.0 / f for f in fresnel]
    return (Propagation(n, nz, ntheta, len(dist), 1.0, 1.0, dist),
            orc.Propagation(n, nz, ntheta, len(dist), 1.0, 1.0, dist))


@pytest.mark.parametrize("n,nz,B", [(8, 8, 2), (16, 32, 3), (64, 64, 2), (256, 128, 2), (32, 1024, 1),
                                    (2048, 16, 1)])
def test_D_DT_vs_oracle_white_noise(gpu, n, nz, B):
    print(f"[case] D/DT white noise n={n} nz={nz} B={B}", flush=True)
    rng = np.random.default_rng(1000 + n + nz)
    g, o = make(n, nz, B, [1e-3, 2e-3])
    x = crandn(rng, (B, nz, n))
    xt = torch.from_numpy(x).to(gpu)
    for j in range(2):
        d = g.D(xt, j).cpu().numpy()
        dt = g.DT(xt, j).cpu().numpy()
        assert rel(d, o.D(x, j)) < REL_TOL
        assert rel(dt, o.DT(x, j)) < REL_TOL


def test_adjoint(gpu):
    print("[case] D/DT adjoint n=nz=256", flush=True)
    rng = np.random.default_rng(7)
    g, _ = make(256, 256, 4, [1e-3, 1.2e-3, 1.5e-3, 2e-3])
    for j in range(4):
        psi = torch.from_numpy(crandn(rng, (4, 256, 256))).to(gpu)
        phi = torch.from_numpy(crandn(rng, (4, 256, 256))).to(gpu)
        lhs = torch.sum(g.D(psi, j).to(torch.complex128) * phi.to(torch.complex128).conj()).item()
        rhs = torch.sum(psi.to(torch.complex128) * g.DT(phi, j).to(torch.complex128).conj()).item()
        assert abs(lhs - rhs) / abs(lhs) < REL_TOL


def test_shapes_and_inplace_semantics(gpu):
    print("[case] D/DT shapes", flush=True)
    g, _ = make(64, 32, 4, [1e-3])
    p2 = torch.ones(32, 64, dtype=torch.complex64, device=gpu)
    p3 = torch.ones(4, 32, 64, dtype=torch.complex64, device=gpu)
    assert g.D(p2, 0).shape == (32, 64)
    assert g.D(p3, 0).shape == (4, 32, 64)
    assert g.DT(p2, 0).shape == (32, 64)
    assert g.DT(p3, 0).shape == (4, 32, 64)
    # D(const) = const (the symmetric extension of a constant is constant)
    assert torch.allclose(g.D(p3, 0), p3, atol=2e-6)
    # input is not modified (D returns a new array)
    x = torch.randn(4, 32, 64, dtype=torch.complex64, device=gpu)
    x0 = x.clone()
    g.D(x, 0)
    assert torch.equal(x, x0)
    with pytest.raises(ValueError):
        g.D(torch.ones(4, 32, 32, dtype=torch.complex64, device=gpu), 0)


def c1_phantom(rng, B, n):
    """Sum of 20 random ellipses, amplitude U[0,1] rad; object exp(i*ph - 0.1*ph)."""
    yy, xx = np.mgrid[0:n, 0:n].astype(np.float64) - n / 2
    out = np.empty((B, n, n), np.complex64)
    for b in range(B):
        ph = np.zeros((n, n))
        for _ in range(20):
            cy, cx = rng.uniform(-0.3, 0.3, 2) * n
            ay, ax = rng.uniform(0.03, 0.2, 2) * n
            t = rng.uniform(0, np.pi)
            u = (xx - cx) * np.cos(t) + (yy - cy) * np.sin(t)
            v = -(xx - cx) * np.sin(t) + (yy - cy) * np.cos(t)
            ph += rng.uniform(0, 1) * ((u / ax) ** 2 + (v / ay) ** 2 <= 1)
        out[b] = np.exp(1j * ph - 0.1 * ph)
    return out


def test_c1_config_D_then_DT(gpu):
    """Batch 16 of 256^2 fields (padded grid 512^2), Fresnel numbers {1e-3,1.2e-3,1.5e-3,2e-3}."""
    print("[case] C1 D->DT 16x256^2 x 4 distances", flush=True)
    rng = np.random.default_rng(1001)
    fres = [1e-3, 1.2e-3, 1.5e-3, 2e-3]
    g, o = make(256, 256, 16, fres)
    x = c1_phantom(rng, 16, 256)
    xt = torch.from_numpy(x).to(gpu)
    for j in range(4):
        y = g.D(xt, j)
        z = g.DT(y, j).cpu().numpy()
        yo = o.D(x.astype(np.complex128), j)
        zo = o.DT(yo, j)
        assert rel(y.cpu().numpy(), yo) < REL_TOL
        assert rel(z, zo) < REL_TOL
