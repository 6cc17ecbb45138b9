"""GPU parity of the back-shift (Shift.Sback, coeffback, curlySback) against
the FP64 oracle: relative L2 <= 1e-5."""
import numpy as np
import pytest
import torch

from oracle import shift as osh

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b))


def crandn(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.mark.parametrize("geom", [(32, 32, 32, 32, 2), (24, 64, 20, 32, 3), (64, 128, 64, 64, 1)])
@pytest.mark.parametrize("m", [1.0, 0.6, 0.85, 1.4])
def test_sback_matches_oracle(gpu, geom, m):
    from holotomocupy_amd import Shift
    n, npsi, nz, nzpsi, B = geom
    print(f"[case] sback geom={geom} m={m}", flush=True)
    rng = np.random.default_rng(n + npsi)
    sh = Shift(n, npsi, nz, nzpsi, "complex64")
    g = crandn(rng, (B, nz, n))
    r = rng.uniform(-3, 3, (B, 2)).astype(np.float32)
    mm = np.full(B, m, dtype=np.float32)
    got = sh.Sback(torch.from_numpy(g).cuda(), torch.from_numpy(r).cuda(), torch.from_numpy(mm).cuda())
    ref = osh.sback(g.astype(np.complex128), r.astype(np.float64), mm.astype(np.float64), n, npsi, nz, nzpsi)
    assert rel(got.cpu().numpy(), ref) < 1e-5


@pytest.mark.parametrize("m", [1.0, 0.7, 1.3])
def test_coeffback_and_curlySback(gpu, m):
    from holotomocupy_amd import Shift
    print(f"[case] coeffback m={m}", flush=True)
    rng = np.random.default_rng(9)
    n, npsi, nz, nzpsi = 32, 64, 32, 32
    sh = Shift(n, npsi, nz, nzpsi, "complex64")
    psi = crandn(rng, (2, nzpsi, npsi))
    got = sh.coeffback(torch.from_numpy(psi).cuda(), m)
    ref = osh.coeffback(psi.astype(np.complex128), np.float32(m).astype(np.float64))
    assert rel(got.cpu().numpy(), ref) < 1e-5
    g = crandn(rng, (1, nz, n))
    r = rng.uniform(-2, 2, (1, 2)).astype(np.float32)
    got2 = sh.curlySback(torch.from_numpy(g).cuda(), torch.from_numpy(r).cuda(), np.float32(m))
    ref2 = osh.coeffback(osh.sback(g.astype(np.complex128), r.astype(np.float64), np.array([np.float32(m)], dtype=np.float64),
                                   n, npsi, nz, nzpsi), np.float32(m).astype(np.float64))
    assert rel(got2.cpu().numpy(), ref2) < 1e-5
