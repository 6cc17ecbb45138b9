"""Sizes that are not powers of two (and powers of two above the limits of
the fixed-size kernels) against the FP64 oracle.

The reference takes any size through cuFFT plans (propagation.py:20-31, 44;
tomo.py:49-54), and its own reconstructions use such sizes: nz = n = 1536
(experimental/y350a_80um/config_step6.conf:11-14), object grids of 1608 and
3264 (experimental/Y350a_dist1234/config_step6.conf:16,
experimental/performance_tests/config1.conf:9-12).  Here these sizes run on
the run-time mixed-radix LDS FFT (csrc/fft_any.hpp).

Bars as for the power-of-two kernels: relative L2 <= 1e-5 per operator and
<= 1e-5 for the adjoint dot-product tests.
"""
import numpy as np
import pytest
import torch

from oracle import propagation as op
from oracle import tomo as ot

pytestmark = pytest.mark.gpu
REL_TOL = 1e-5


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.linalg.norm(a.astype(np.complex128) - b) / np.linalg.norm(b))


def crandn(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.fixture(autouse=True)
def _workers():
    op.WORKERS = 16


@pytest.mark.parametrize("n,nz,B", [(12, 10, 2), (24, 40, 2), (100, 60, 2), (96, 1000, 1), (1536, 1536, 1),
                                    (1536, 1608, 1), (2500, 2500, 1), (4096, 8, 1), (8, 4096, 1)])
def test_D_DT_any_size_vs_oracle(gpu, n, nz, B):
    from holotomocupy_amd import Propagation
    print(f"[case] D/DT any size n={n} nz={nz} B={B}", flush=True)
    rng = np.random.default_rng(2000 + n + nz)
    dist = [1.0 / 1e-3, 1.0 / 2e-3]
    g = Propagation(n, nz, B, 2, 1.0, 1.0, dist)
    o = op.Propagation(n, nz, B, 2, 1.0, 1.0, dist)
    x = crandn(rng, (B, nz, n))
    xt = torch.from_numpy(x).to(gpu)
    for j in range(2):
        assert rel(g.D(xt, j), o.D(x.astype(np.complex128), j)) < REL_TOL, ("D", j)
        assert rel(g.DT(xt, j), o.DT(x.astype(np.complex128), j)) < REL_TOL, ("DT", j)


def test_D_DT_1536_bench_physics_adjoint(gpu):
    """nz = n = 1536 at the physics of a 1536^2 detector (17.1 keV, 4 distances)."""
    from holotomocupy_amd import Propagation
    from holotomocupy_amd.rec import physics
    print("[case] D/DT 1536^2 bench physics, adjoint", flush=True)
    z1 = np.array([5.110, 5.464, 6.879, 9.817]) * 1e-3
    ph = physics(17.1, z1, 1.217, 1.4760147601476e-6 * 2048 / 1536)
    g = Propagation(1536, 1536, 2, 4, ph.wavelength, ph.voxelsize, ph.distance)
    rng = np.random.default_rng(2101)
    for j in (0, 3):
        psi = torch.from_numpy(crandn(rng, (2, 1536, 1536))).to(gpu)
        phi = torch.from_numpy(crandn(rng, (2, 1536, 1536))).to(gpu)
        lhs = torch.sum(g.D(psi, j).to(torch.complex128) * phi.to(torch.complex128).conj()).item()
        rhs = torch.sum(psi.to(torch.complex128) * g.DT(phi, j).to(torch.complex128).conj()).item()
        assert abs(lhs - rhs) / abs(lhs) < REL_TOL
    o = op.Propagation(1536, 1536, 1, 4, ph.wavelength, ph.voxelsize, ph.distance)
    x = crandn(rng, (1, 1536, 1536))
    assert rel(g.D(torch.from_numpy(x).to(gpu), 2), o.D(x.astype(np.complex128), 2)) < REL_TOL


def test_odd_sizes_rejected(gpu):
    from holotomocupy_amd import Propagation, Tomo
    from holotomocupy_amd._lib import HtamdError
    with pytest.raises(HtamdError, match="even"):
        Propagation(13, 16, 1, 1, 1.0, 1.0, [1.0])
    with pytest.raises(HtamdError, match="even"):
        Tomo(13, 2, np.linspace(0, np.pi, 8, endpoint=False), 0)


@pytest.mark.parametrize("n,nth,nz,mask_r", [(12, 10, 2, 0.0), (40, 24, 3, 0.9), (50, 30, 2, 0.9), (96, 50, 2, 0.9),
                                             (1608, 90, 2, 0.0), (2048, 90, 2, 0.9), (2500, 60, 2, 0.0)])
def test_R_RT_any_size_vs_oracle(gpu, n, nth, nz, mask_r):
    from holotomocupy_amd import Tomo
    print(f"[case] R/RT any size n={n} ntheta={nth} nz={nz} mask_r={mask_r}", flush=True)
    theta = np.arange(nth) * np.pi / nth
    rng = np.random.default_rng(2200 + n)
    obj = crandn(rng, (nz, n, n))
    sino = crandn(rng, (nth, nz, n))
    g, o = Tomo(n, nz, theta, mask_r), ot.Tomo(n, nz, theta, mask_r)
    Rg = g.R(torch.from_numpy(obj).to(gpu))
    RTg = g.RT(torch.from_numpy(sino).to(gpu))
    assert rel(Rg, o.R(obj.astype(np.complex128))) < REL_TOL
    assert rel(RTg, o.RT(sino.astype(np.complex128))) < REL_TOL
    lhs = np.vdot(Rg.cpu().numpy().astype(np.complex128), sino)
    rhs = np.vdot(obj.astype(np.complex128), RTg.cpu().numpy())
    assert abs(lhs - rhs) / abs(lhs) < REL_TOL
    # float32 slices take the real path
    objr = rng.standard_normal((nz, n, n)).astype(np.float32)
    assert rel(g.R(torch.from_numpy(objr).to(gpu)), o.R(objr.astype(np.float64)).real) < REL_TOL


@pytest.mark.parametrize("filter_name", ["ramp", "shepp", "parzen"])
def test_filter_and_fbp_any_size(gpu, filter_name):
    from holotomocupy_amd import Tomo
    print(f"[case] filter / fbp any size n=100 filter={filter_name}", flush=True)
    n, nth, nz = 100, 60, 3
    theta = np.arange(nth) * np.pi / nth
    rng = np.random.default_rng(2300)
    sino = rng.standard_normal((nth, nz, n)).astype(np.float32)
    g, o = Tomo(n, nz, theta, 0.0), ot.Tomo(n, nz, theta, 0.0)
    st = torch.from_numpy(sino).to(gpu)
    assert rel(g._filter_sino(st, filter_name), o.filter_sino(sino.astype(np.float64), filter_name).real) < REL_TOL
    assert rel(g.fbp(st, filter_name), o.fbp(sino.astype(np.float64), filter_name).real) < REL_TOL


def test_bh_loop_any_size_matches_oracle(gpu):
    """Three BH iterations at a 48^2 detector and a 24^3 object: the fused
    passes take power-of-two detectors only, so the loop runs its gradient
    and bilinear forms on the per-operator kernels (mixed-radix D/DT, R/RT of
    n = 24), against the FP64 restatement (bars as the power-of-two loop
    test: 1e-4)."""
    from test_gpu_bhloop import make_bh
    print("[case] BH loop n=48 nobj=24 (per-operator path)", flush=True)
    rec, orc, start = make_bh(False, 0.3, 0.05, n=48, nobj=24, ntheta=6, nchunk=3)
    assert not rec.fused_available()
    rec.BH()
    torch.cuda.synchronize()
    vo, hist = orc.BH(start, rec.niter)
    errs = [e for i, e, _ in rec.table if i >= 0]
    for k in range(rec.niter):
        assert float(rec.alphas[k]) == pytest.approx(hist[k][0], rel=1e-4)
        assert errs[k] == pytest.approx(hist[k][2], rel=1e-4)
    for k in ("obj", "prb", "pos"):
        assert rel(rec.vars[k], vo[k]) < 1e-4, k
