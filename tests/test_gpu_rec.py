"""GPU parity of the BH cascade (holotomocupy_amd.rec.Rec) against the FP64
oracle cascade.  Tolerances: relative L2 <= 1e-5 for fields and the object /
probe gradients, 1e-4 for the position gradient and the scalar bilinear forms
(sums of many signed terms, FP32 accumulation inside the kernels)."""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from oracle import cascade as oc
from oracle import propagation as op
from oracle import shift as osh

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def make_problem(real=False, n=32, nobj=40, ndist=3, B=4, seed=5, shrink=None):
    from holotomocupy_amd.rec import Rec, physics
    rng = np.random.default_rng(seed)
    args = SimpleNamespace(energy=17.1, detector_pixelsize=1.4760147601476e-6 * 16,
                           focustodetectordistance=1.217,
                           z1=np.array([5.110, 5.464, 6.879, 9.817])[:ndist] * 1e-3,
                           ndist=ndist, ntheta=B, nz=n, n=n, nzobj=nobj, nobj=nobj,
                           obj_dtype="float32" if real else "complex64", rho=[1, 0.5, 0.2])
    if shrink is not None:  # per-angle, per-distance magnification change (rec_mpi.py:146)
        args.shrink = shrink
    rec = Rec(args)
    ph = physics(args.energy, args.z1, args.focustodetectordistance, args.detector_pixelsize)
    prop = op.Propagation(n, n, B, ndist, ph.wavelength, ph.voxelsize, ph.distance)
    sh = osh.Shift(n, nobj, n, nobj, args.obj_dtype)
    mag = np.tile(1 / ph.norm_magnifications, (B, 1))
    if shrink is not None:
        mag = (1 + np.asarray(shrink)) / ph.norm_magnifications[None, :]
    cas = oc.Cascade(prop, sh, ndist, n, n, nobj, nobj, B * ndist * n * n, mag, rho=args.rho, obj_real=real)
    yy, xx = np.mgrid[0:n, 0:n] - n / 2
    prb = np.stack([np.exp(-(xx ** 2 + yy ** 2) / (2 * (0.4 * n) ** 2)) * np.exp(0.3j * np.cos(0.2 * xx + j))
                    for j in range(ndist)])
    prb = (prb * (1 + 0.05 * crandn(rng, prb.shape))).astype(np.complex64)
    proj = 0.3 * (rng.standard_normal((B, nobj, nobj)) if real else crandn(rng, (B, nobj, nobj)))
    proj = proj.astype(np.float32 if real else np.complex64)
    pos = rng.uniform(-3, 3, (B, ndist, 2)).astype(np.float32)
    d = (np.abs(cas.forward(prb, proj, pos)) * (1 + 0.05 * rng.standard_normal((B, ndist, n, n)))).astype(np.float32)
    x = {"prb": (prb * (1 + 0.1 * crandn(rng, prb.shape))).astype(np.complex64),
         "proj": (proj + 0.05 * (rng.standard_normal(proj.shape) if real else crandn(rng, proj.shape))).astype(proj.dtype),
         "pos": (pos + rng.uniform(-0.5, 0.5, pos.shape)).astype(np.float32)}
    v = {"prb": (0.1 * crandn(rng, prb.shape)).astype(np.complex64),
         "proj": (0.1 * (rng.standard_normal(proj.shape) if real else crandn(rng, proj.shape))).astype(proj.dtype),
         "pos": (0.3 * rng.standard_normal(pos.shape)).astype(np.float32)}
    w = {"prb": (0.1 * crandn(rng, prb.shape)).astype(np.complex64),
         "proj": (0.1 * (rng.standard_normal(proj.shape) if real else crandn(rng, proj.shape))).astype(proj.dtype),
         "pos": (0.3 * rng.standard_normal(pos.shape)).astype(np.float32)}
    return rec, cas, x, v, w, d


def dev(dct):
    return {k: torch.from_numpy(np.ascontiguousarray(a)).cuda() for k, a in dct.items()}


def lst(dct):
    return [dct["prb"].astype(np.complex128), dct["proj"].astype(np.float64 if dct["proj"].dtype == np.float32 else np.complex128),
            dct["pos"].astype(np.float64)]


@pytest.mark.parametrize("real", [False, True])
def test_forward_objective_gradient(gpu, real):
    print(f"[case] cascade forward/gradient real={real}", flush=True)
    rec, cas, x, v, w, d = make_problem(real)
    xd, dd = dev(x), torch.from_numpy(d).cuda()
    fwd = rec.forward(xd)
    assert rel(fwd, cas.forward(*lst(x))) < 1e-5
    assert rec.objective(xd, dd) == pytest.approx(cas.objective(*lst(x), d), rel=1e-5)
    grads = {"prb": torch.zeros_like(xd["prb"]), "proj": torch.empty_like(xd["proj"]),
             "pos": torch.empty_like(xd["pos"])}
    rec.gradients_cascade(xd, grads, dd)
    gprb, gproj, gpos = cas.gradients(*lst(x), d)
    assert rel(grads["prb"], gprb) < 1e-5
    assert rel(grads["proj"], gproj) < 1e-5
    assert rel(grads["pos"], gpos) < 1e-4


@pytest.mark.parametrize("real", [False, True])
def test_hessian(gpu, real):
    print(f"[case] cascade hessian real={real}", flush=True)
    rec, cas, x, v, w, d = make_problem(real, seed=6)
    xd, vd, wd, dd = dev(x), dev(v), dev(w), torch.from_numpy(d).cuda()
    h_vv = rec.hessian_cascade(xd, vd, vd, dd)
    h_vw = rec.hessian_cascade(xd, vd, wd, dd)
    lx, lv, lw = lst(x), lst(v), lst(w)
    assert h_vv == pytest.approx(cas.hessian(lx, lv, lv, d), rel=1e-4)
    assert h_vw == pytest.approx(cas.hessian(lx, lv, lw, d), rel=1e-4)
