"""Per-operator error against the FP64 oracle at the working sizes: D/DT at
2048^2 (the holotomography physics of the bench, 4 distances), R/RT at 256^2
with 384 angles (a 4-slice slab), coeff/S/Sadj from a 1024^2 object onto a
2048^2 detector (magnifications 1 and 1.85), and the BH objective, gradient
and bilinear form at 2048^2 / 1024^2 / 4 distances on one angle.  Each case
prints its relative L2 error; the bars are those of the other parity tests
(1e-5 per operator and for the probe and object gradients, 1e-4 for the
position gradient, the objective and the bilinear form, which are sums of
many signed FP32 terms)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

Z1 = np.array([5.110, 5.464, 6.879, 9.817]) * 1e-3


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.linalg.norm(a.astype(np.complex128) - b) / np.linalg.norm(b))


def report(name, err, bar):
    print(f"[precision] {name}: {err:.2e} (bar {bar:.0e})", flush=True)
    assert err < bar, name


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_propagator_2048(gpu):
    from holotomocupy_amd import Propagation, phantom
    from holotomocupy_amd.rec import physics
    from oracle import propagation as op
    print("[case] precision D/DT 2048^2", flush=True)
    op.WORKERS = 16
    n = 2048
    ph = physics(17.1, Z1, 1.217, 1.4760147601476e-6 * 2)
    g = Propagation(n, n, 1, 4, ph.wavelength, ph.voxelsize, ph.distance)
    o = op.Propagation(n, n, 1, 4, ph.wavelength, ph.voxelsize, ph.distance)
    rng = np.random.default_rng(41)
    smooth = phantom.probe(rng, 1, n, n)
    white = crandn(rng, (1, n, n)).astype(np.complex64)
    report("D_0 smooth 2048^2", rel(g.D(torch.from_numpy(smooth).cuda(), 0), o.D(smooth.astype(np.complex128), 0)),
           1e-5)
    report("D_3 white 2048^2", rel(g.D(torch.from_numpy(white).cuda(), 3), o.D(white.astype(np.complex128), 3)),
           1e-5)
    report("DT_1 white 2048^2", rel(g.DT(torch.from_numpy(white).cuda(), 1), o.DT(white.astype(np.complex128), 1)),
           1e-5)


def test_radon_256_a384(gpu):
    from holotomocupy_amd import Tomo
    from oracle import tomo as ot
    print("[case] precision R/RT 256^2, 384 angles, 4 slices", flush=True)
    n, nth, nz = 256, 384, 4
    theta = np.arange(nth) * np.pi / nth
    rng = np.random.default_rng(42)
    obj = crandn(rng, (nz, n, n)).astype(np.complex64)
    sino = crandn(rng, (nth, nz, n)).astype(np.complex64)
    g, o = Tomo(n, nz, theta, 0.9), ot.Tomo(n, nz, theta, 0.9)
    report("R 256^2 a384", rel(g.R(torch.from_numpy(obj).cuda()), o.R(obj.astype(np.complex128))), 1e-5)
    report("RT 256^2 a384", rel(g.RT(torch.from_numpy(sino).cuda()), o.RT(sino.astype(np.complex128))), 1e-5)


def test_bspline_2048(gpu):
    from holotomocupy_amd import Shift, phantom
    from oracle import shift as osh
    print("[case] precision coeff/S/Sadj 1024^2 object, 2048^2 detector", flush=True)
    osh.WORKERS = 16
    n, npsi = 2048, 1024
    rng = np.random.default_rng(43)
    c = phantom.projections(rng, 2, npsi, npsi)
    s = crandn(rng, (2, n, n)).astype(np.complex64)
    r = rng.uniform(-3, 3, (2, 2)).astype(np.float32)
    m = np.array([1.0, 1.85], dtype=np.float32)
    g, o = Shift(n, npsi, n, npsi, "complex64"), osh.Shift(n, npsi, n, npsi, "complex64")
    rd, md = torch.from_numpy(r).cuda(), torch.from_numpy(m).cuda()
    report("coeff 1024^2", rel(g.coeff(torch.from_numpy(c).cuda()), o.coeff(c.astype(np.complex128))), 1e-5)
    report("S 1024^2 -> 2048^2", rel(g.S(torch.from_numpy(c).cuda(), rd, md), o.S(c.astype(np.complex128), r, m)),
           1e-5)
    report("Sadj 2048^2 -> 1024^2", rel(g.Sadj(torch.from_numpy(s).cuda(), rd, md),
                                        o.Sadj(s.astype(np.complex128), r, m)), 1e-5)


def test_bh_cascade_2048(gpu):
    """One angle at the bench's shapes, evaluated away from the true point
    (data from perturbed variables, 1% noise)."""
    from types import SimpleNamespace
    from holotomocupy_amd import phantom
    from holotomocupy_amd.rec import Rec, physics
    from oracle import cascade as oc
    from oracle import propagation as op
    from oracle import shift as osh
    print("[case] precision BH objective / gradient / bilinear form 2048^2, 1 angle", flush=True)
    op.WORKERS = osh.WORKERS = 16
    n, nobj, nd = 2048, 1024, 4
    pix = 1.4760147601476e-6 * 2
    args = SimpleNamespace(energy=17.1, detector_pixelsize=pix, focustodetectordistance=1.217, z1=Z1, ndist=nd,
                           ntheta=1, nz=n, n=n, nzobj=nobj, nobj=nobj, obj_dtype="complex64", rho=[1.0, 0.05, 0.02])
    rec = Rec(args)
    ph = physics(17.1, Z1, 1.217, pix)
    cas = oc.Cascade(op.Propagation(n, n, 1, nd, ph.wavelength, ph.voxelsize, ph.distance),
                     osh.Shift(n, nobj, n, nobj), nd, n, n, nobj, nobj, nd * n * n,
                     (1 / ph.norm_magnifications)[None, :], rho=(1.0, 0.05, 0.02))
    rng = np.random.default_rng(44)
    prb = phantom.probe(rng, nd, n, n)
    proj = phantom.projections(rng, 1, nobj, nobj)
    pos = phantom.positions(rng, 1, nd)
    amp = np.abs(cas.forward(prb.astype(np.complex128), proj.astype(np.complex128), pos.astype(np.float64)))
    d = (amp * (1 + 0.01 * rng.standard_normal(amp.shape))).astype(np.float32)
    x = [(prb * 0.9).astype(np.complex64), (proj * 0.8).astype(np.complex64),
         (pos + rng.uniform(-0.3, 0.3, pos.shape)).astype(np.float32)]
    y = [(0.05 * crandn(rng, prb.shape)).astype(np.complex64), (0.05 * crandn(rng, proj.shape)).astype(np.complex64),
         (0.3 * rng.standard_normal(pos.shape)).astype(np.float32)]
    wide = [x[0].astype(np.complex128), x[1].astype(np.complex128), x[2].astype(np.float64)]
    ywide = [y[0].astype(np.complex128), y[1].astype(np.complex128), y[2].astype(np.float64)]
    dev = lambda v: torch.from_numpy(np.ascontiguousarray(v)).cuda()  # noqa: E731
    xd = {"prb": dev(x[0]), "proj": dev(x[1]), "pos": dev(x[2])}
    yd = {"prb": dev(y[0]), "proj": dev(y[1]), "pos": dev(y[2])}
    dd = dev(d)
    d64 = d.astype(np.float64)
    obj = rec.objective(xd, dd)
    report("BH objective", abs(obj - cas.objective(*wide, d64)) / abs(cas.objective(*wide, d64)), 1e-4)
    g = {"prb": torch.zeros_like(xd["prb"]), "proj": torch.empty_like(xd["proj"]), "pos": torch.empty_like(xd["pos"])}
    rec.gradients_fused(xd, g, dd)
    gprb, gproj, gpos = cas.gradients(*wide, d64)
    report("gradient probe (fused)", rel(g["prb"], gprb), 1e-5)
    report("gradient object (fused)", rel(g["proj"], gproj), 1e-5)
    report("gradient positions (fused)", rel(g["pos"], gpos), 1e-4)
    h = float(rec.hessian_fused(xd, yd, yd, dd))
    href = cas.hessian(wide, ywide, ywide, d64)
    report("bilinear form H(y, y) (fused)", abs(h - href) / abs(href), 1e-4)
