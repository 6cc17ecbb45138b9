"""Parity at the full forward + BH gradient configuration (object 512^3,
720 angles, 4 distances, detector and probe 1024^2): the instantiations of
the fused gradient and bilinear-form passes for 1024-point transforms and of
the Radon kernels at n = 512, against the FP64 oracle on the same inputs.

Inputs follow that configuration: Gaussian probe with smooth random phase,
ellipse-phantom projections on a 512^2 object plane, shifts U[-3, 3] px per
angle and distance, and float32 amplitudes sqrt(Poisson(1e4 |F|^2) / 1e4);
the physics is the bench's (17.1 keV, four focus-to-sample distances,
1.476 um * 4 pixels at 1024^2).  The point is perturbed from the truth so the
residual is not just the noise.

Bars: 1e-5 relative L2 for the probe and object gradients and for R / RT;
1e-4 for the position gradient, the objective and the bilinear forms (sums of
many signed FP32 terms).  The fused gradient is also checked against central
differences of the FP64 objective (bar 1e-4: the step's O(h^2) truncation is
~1e-6, the rest is the gradient's FP32 error)."""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

Z1 = np.array([5.110, 5.464, 6.879, 9.817]) * 1e-3
N, NOBJ, ND = 1024, 512, 4
PIX = 1.4760147601476e-6 * 2 * (2048 // N)
RHO = (1.0, 0.05, 0.02)


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.linalg.norm(a.astype(np.complex128) - b) / np.linalg.norm(b))


def report(name, err, bar):
    print(f"[c3] {name}: {err:.2e} (bar {bar:.0e})", flush=True)
    assert err < bar, name


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _problem(B, seed):
    from holotomocupy_amd import phantom
    from holotomocupy_amd.rec import Rec, physics
    from oracle import cascade as oc
    from oracle import propagation as op
    from oracle import shift as osh
    op.WORKERS = osh.WORKERS = 16
    args = SimpleNamespace(energy=17.1, detector_pixelsize=PIX, focustodetectordistance=1.217, z1=Z1, ndist=ND,
                           ntheta=B, nz=N, n=N, nzobj=NOBJ, nobj=NOBJ, obj_dtype="complex64", rho=list(RHO))
    rec = Rec(args)
    ph = physics(17.1, Z1, 1.217, PIX)
    cas = oc.Cascade(op.Propagation(N, N, B, ND, ph.wavelength, ph.voxelsize, ph.distance),
                     osh.Shift(N, NOBJ, N, NOBJ), ND, N, N, NOBJ, NOBJ, B * ND * N * N,
                     np.tile(1 / ph.norm_magnifications, (B, 1)), rho=RHO)
    rng = np.random.default_rng(seed)
    prb = phantom.probe(rng, ND, N, N)
    proj = phantom.projections(rng, B, NOBJ, NOBJ)
    pos = phantom.positions(rng, B, ND, 3.0)
    amp = np.abs(cas.forward(prb.astype(np.complex128), proj.astype(np.complex128), pos.astype(np.float64)))
    d = np.sqrt(rng.poisson(1e4 * amp ** 2) / 1e4).astype(np.float32)
    x = [(prb * 0.9).astype(np.complex64), (proj * 0.8).astype(np.complex64),
         (pos + rng.uniform(-0.3, 0.3, pos.shape)).astype(np.float32)]
    return rec, cas, x, d, rng


def _wide(v):
    return [v[0].astype(np.complex128), v[1].astype(np.complex128), v[2].astype(np.float64)]


def _dev(v):
    return {k: torch.from_numpy(np.ascontiguousarray(a)).cuda() for k, a in zip(("prb", "proj", "pos"), v)}


def test_c3_fused_gradient_objective_and_differences(gpu):
    print("[case] C3 fused gradient 1024^2 / 512^2 / 4 distances, 2 angles", flush=True)
    B = 2
    rec, cas, x, d, rng = _problem(B, 1003)
    xd, dd = _dev(x), torch.from_numpy(d).cuda()
    d64 = d.astype(np.float64)
    ref_obj = cas.objective(*_wide(x), d64)
    report("objective", abs(rec.objective(xd, dd) - ref_obj) / abs(ref_obj), 1e-4)
    gprb, gproj, gpos = cas.gradients(*_wide(x), d64)
    for path in ("gradients_fused", "gradients_cascade"):
        g = {"prb": torch.zeros_like(xd["prb"]), "proj": torch.empty_like(xd["proj"]),
             "pos": torch.empty_like(xd["pos"])}
        getattr(rec, path)(xd, g, dd)
        report(f"{path} probe", rel(g["prb"], gprb), 1e-5)
        report(f"{path} object", rel(g["proj"], gproj), 1e-5)
        report(f"{path} positions", rel(g["pos"], gpos), 1e-4)
        if path == "gradients_fused":
            gf = g
    # central differences of the FP64 objective along a random direction,
    # (f(x + h v) - f(x - h v)) / 2h = sum_k Re<grad_k / rho_k^2, v_k> + O(h^2):
    # at 1e-3 of the variables the truncation is ~1e-6 relative, so what the
    # check sees is the GPU gradient's own error
    v = [(1e-3 * crandn(rng, x[0].shape) * np.abs(x[0]).mean()),
         (1e-3 * crandn(rng, x[1].shape) * np.abs(x[1]).mean()),
         (5e-4 * rng.standard_normal(x[2].shape))]
    xw = _wide(x)

    def f(s):
        return cas.objective(*[a + s * b for a, b in zip(xw, v)], d64)

    fd = (f(1.0) - f(-1.0)) / 2.0
    dd_an = 0.0
    # rho is (object, probe, positions)
    for k, vk, r in zip(("prb", "proj", "pos"), v, (RHO[1], RHO[0], RHO[2])):
        gk = gf[k].cpu().numpy()
        dd_an += float(np.sum(np.real(np.conj(gk.astype(np.complex128)) * vk))) / r ** 2
    report("central differences (FP64 objective) vs GPU gradient", abs(fd - dd_an) / abs(dd_an), 1e-4)


def test_c3_fused_bilinear_forms(gpu):
    print("[case] C3 fused bilinear forms (general, equal, pair) 1024^2 / 512^2, 2 angles", flush=True)
    B = 2
    rec, cas, x, d, rng = _problem(B, 1013)
    y = [(0.05 * crandn(rng, x[0].shape)).astype(np.complex64), (0.05 * crandn(rng, x[1].shape)).astype(np.complex64),
         (0.3 * rng.standard_normal(x[2].shape)).astype(np.float32)]
    z = [(0.03 * crandn(rng, x[0].shape)).astype(np.complex64), (0.07 * crandn(rng, x[1].shape)).astype(np.complex64),
         (0.2 * rng.standard_normal(x[2].shape)).astype(np.float32)]
    xd, yd, zd, dd = _dev(x), _dev(y), _dev(z), torch.from_numpy(d).cuda()
    d64 = d.astype(np.float64)
    h_yz = cas.hessian(_wide(x), _wide(y), _wide(z), d64)
    h_zz = cas.hessian(_wide(x), _wide(z), _wide(z), d64)
    h_yy = cas.hessian(_wide(x), _wide(y), _wide(y), d64)
    report("H(y, z) general", abs(float(rec.hessian_fused(xd, yd, zd, dd)) - h_yz) / abs(h_yz), 1e-4)
    report("H(y, y) equal", abs(float(rec.hessian_fused(xd, yd, yd, dd)) - h_yy) / abs(h_yy), 1e-4)
    pair = rec.hessian_pair_fused(xd, yd, zd, dd).cpu().numpy()
    report("pair H(y, z)", abs(pair[0] - h_yz) / abs(h_yz), 1e-4)
    report("pair H(z, z)", abs(pair[1] - h_zz) / abs(h_zz), 1e-4)
    report("H(y, z) per-operator", abs(rec.hessian_cascade(xd, yd, zd, dd) - h_yz) / abs(h_yz), 1e-4)


@pytest.mark.parametrize("mask_r", [0.0, 0.9])
def test_c3_radon_512_a720(gpu, mask_r):
    from holotomocupy_amd import Tomo
    from oracle import tomo as ot
    print(f"[case] C3 R/RT 512^2, 720 angles, 3 slices, mask_r={mask_r}", flush=True)
    n, nth, nz = 512, 720, 3
    theta = np.arange(nth) * np.pi / nth
    rng = np.random.default_rng(1023)
    obj = crandn(rng, (nz, n, n)).astype(np.complex64)
    sino = crandn(rng, (nth, nz, n)).astype(np.complex64)
    g, o = Tomo(n, nz, theta, mask_r), ot.Tomo(n, nz, theta, mask_r)
    Rg = g.R(torch.from_numpy(obj).cuda())
    RTg = g.RT(torch.from_numpy(sino).cuda())
    report("R 512^2 a720", rel(Rg, o.R(obj.astype(np.complex128))), 1e-5)
    report("RT 512^2 a720", rel(RTg, o.RT(sino.astype(np.complex128))), 1e-5)
    # adjoint dot-product test on the GPU results
    lhs = np.vdot(Rg.cpu().numpy().astype(np.complex128), sino)
    rhs = np.vdot(obj.astype(np.complex128), RTg.cpu().numpy())
    report("adjoint <R x, y> vs <x, RT y>", abs(lhs - rhs) / abs(lhs), 1e-5)
    # a float32 object takes the real path
    objr = rng.standard_normal((nz, n, n)).astype(np.float32)
    report("R 512^2 a720 float32", rel(g.R(torch.from_numpy(objr).cuda()), o.R(objr.astype(np.float64)).real), 1e-5)
