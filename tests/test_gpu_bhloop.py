"""GPU parity of the on-device BH loop (holotomocupy_amd.bhloop, Rec.BH) and
its kernels against the FP64 oracle (oracle/bh.py).

Tolerances: the Laplacian kernel and the device-coefficient BLAS-1 forms are
checked at 1e-6 relative (single FP32 stencil sums; FP64 accumulation).  The
loop runs three iterations in FP32 against the FP64 restatement; its step
sizes and objective values are compared at 1e-4 relative and the variables
at 1e-4 relative L2 (the FP32 gradient and bilinear forms are 1e-5 accurate,
and three steps compound that error)."""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from oracle import bh as ob

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("real", [False, True])
def test_laplacian3d_matches_oracle(gpu, real):
    from holotomocupy_amd import _lib
    print(f"[case] laplacian3d real={real}", flush=True)
    rng = np.random.default_rng(11)
    shape = (5, 6, 12)
    u = rng.standard_normal(shape) if real else crandn(rng, shape)
    u32 = u.astype(np.float32 if real else np.complex64)
    ud = torch.from_numpy(u32).cuda()
    nz = shape[0]

    def lapk(x, z0, nz_out, add=None, scale=1.0):
        out = torch.empty((nz_out,) + tuple(x.shape[1:]), dtype=x.dtype, device="cuda")
        _lib.call("htamd_laplacian3d", _lib.ptr(x), x.shape[0], x.shape[1], x.shape[2], 0 if real else 1, z0,
                  nz_out, _lib.ptr(add) if add is not None else None, float(scale), _lib.ptr(out),
                  _lib.stream_ptr())
        return out

    p = ob.pad2(u32.astype(np.float64 if real else np.complex128))
    lap_ext = ob.lap(p[:-2], p[1:-1], p[2:])  # slices -1 .. nz
    L = lapk(ud, -1, nz + 2)
    torch.cuda.synchronize()
    assert rel(L, lap_ext) < 1e-6
    g = torch.from_numpy((rng.standard_normal(shape) if real else crandn(rng, shape)).astype(u32.dtype)).cuda()
    g0 = g.cpu().numpy().astype(np.float64 if real else np.complex128)
    out = lapk(L, 1, nz, add=g, scale=0.37)
    assert rel(out, g0 + 0.37 * ob.biharmonic(u32.astype(g0.dtype))) < 1e-6
    # in-place add (out aliases add) and the window of the volume's own slices
    _lib.call("htamd_laplacian3d", _lib.ptr(L), L.shape[0], L.shape[1], L.shape[2], 0 if real else 1, 1, nz,
              _lib.ptr(g), 0.37, _lib.ptr(g), _lib.stream_ptr())
    assert rel(g, g0 + 0.37 * ob.biharmonic(u32.astype(g0.dtype))) < 1e-6
    assert rel(lapk(ud, 0, nz), lap_ext[1:-1]) < 1e-6


def test_blas1_device_coefficients(gpu):
    from holotomocupy_amd import _lib
    print("[case] blas1 device coefficients", flush=True)
    rng = np.random.default_rng(12)
    x = torch.from_numpy(crandn(rng, 10007).astype(np.complex64)).cuda()
    y = torch.from_numpy(crandn(rng, 10007).astype(np.complex64)).cuda()
    xf, yf = torch.view_as_real(x), torch.view_as_real(y)
    ab = torch.tensor([0.75, -1.25], dtype=torch.float64, device="cuda")
    o1, o2 = torch.empty_like(xf), torch.empty_like(xf)
    _lib.call("htamd_axpby_dc", _lib.ptr(xf), _lib.ptr(yf), _lib.ptr(ab), _lib.ptr(o1), xf.numel(), _lib.stream_ptr())
    _lib.call("htamd_axpby", _lib.ptr(xf), _lib.ptr(yf), 0.75, -1.25, _lib.ptr(o2), xf.numel(), _lib.stream_ptr())
    assert torch.equal(o1, o2)
    ws = torch.empty(1024, dtype=torch.float64, device="cuda")
    res = torch.tensor([2.0], dtype=torch.float64, device="cuda")
    x2 = xf.clone()
    _lib.call("htamd_axpby_redot_dc", _lib.ptr(x2), _lib.ptr(yf), _lib.ptr(ab), xf.numel(), _lib.ptr(res), -0.5, 1,
              _lib.ptr(ws), ws.numel() * 8, _lib.stream_ptr())
    xn = 0.75 * x.cpu().numpy().astype(np.complex128) - 1.25 * y.cpu().numpy()
    yn = y.cpu().numpy().astype(np.complex128)
    assert torch.equal(x2, o2)
    ref = 2.0 - 0.5 * float(np.sum(xn.real * yn.real + xn.imag * yn.imag))
    assert float(res) == pytest.approx(ref, rel=1e-6)


def make_bh(real=False, lam_prbfit=0.0, lam_laplacian=0.0, n=32, nobj=16, ntheta=8, ndist=2, nchunk=4,
            niter=3, seed=21):
    from holotomocupy_amd.rec import Rec, physics
    rng = np.random.default_rng(seed)
    theta = (np.arange(ntheta) * np.pi / ntheta).astype(np.float32).astype(np.float64)
    args = SimpleNamespace(energy=17.1, detector_pixelsize=1.4760147601476e-6 * 16, focustodetectordistance=1.217,
                           z1=np.array([5.110, 5.464, 6.879, 9.817])[:ndist] * 1e-3, ndist=ndist, ntheta=ntheta,
                           nz=n, n=n, nzobj=nobj, nobj=nobj, obj_dtype="float32" if real else "complex64",
                           rho=[1.0, 0.5, 0.2], theta=theta, mask_r=0, niter=niter, start_iter=0, err_step=1,
                           vis_step=-1, nchunk=nchunk, lam_prbfit=lam_prbfit, lam_laplacian=lam_laplacian)
    rec = Rec(args)
    ph = physics(args.energy, args.z1, args.focustodetectordistance, args.detector_pixelsize)
    mag = np.tile(1 / ph.norm_magnifications, (ntheta, 1))
    orc = ob.BHSolver(n, n, nobj, nobj, ndist, ph.wavelength, ph.voxelsize, ph.distance, theta, mag,
                      rho=args.rho, lam_prbfit=lam_prbfit, lam_laplacian=lam_laplacian, obj_real=real)
    yy, xx = np.mgrid[0:n, 0:n] - n / 2
    prb = np.stack([np.exp(-(xx ** 2 + yy ** 2) / (2 * (0.4 * n) ** 2)) * np.exp(0.3j * np.cos(0.2 * xx + j))
                    for j in range(ndist)]).astype(np.complex64)
    zz, yo, xo = np.mgrid[0:nobj, 0:nobj, 0:nobj] - nobj / 2
    blob = np.exp(-((zz - 1) ** 2 + (yo + 2) ** 2 + (xo - 1) ** 2) / (2 * (0.22 * nobj) ** 2))
    obj = (0.5 * blob if real else 0.5 * blob * (1 + 0.3j)).astype(np.float32 if real else np.complex64)
    pos = rng.uniform(-2, 2, (ntheta, ndist, 2)).astype(np.float32)
    c128 = lambda a: a.astype(np.float64 if a.dtype == np.float32 else np.complex128)  # noqa: E731
    truth = {"obj": c128(obj), "prb": c128(prb), "pos": pos.astype(np.float64)}
    data = orc.gen_sqrt_data(truth).astype(np.float32)
    ref = (orc.gen_sqrt_ref(truth["prb"]) * 1.01).astype(np.float32)
    orc.data, orc.ref = data.astype(np.float64), ref.astype(np.float64)
    noise = rng.standard_normal(obj.shape) if real else crandn(rng, obj.shape)
    start = {"obj": (0.8 * obj + 0.01 * noise).astype(obj.dtype),
             "prb": (prb * (1 + 0.03 * crandn(rng, prb.shape))).astype(np.complex64),
             "pos": (pos + 0.25 * rng.standard_normal(pos.shape)).astype(np.float32)}
    rec.data.copy_(torch.from_numpy(data))
    rec.ref.copy_(torch.from_numpy(ref))
    for k in ("obj", "prb", "pos"):
        rec.vars[k].copy_(torch.from_numpy(start[k]))
    return rec, orc, {k: c128(v) for k, v in start.items()}


@pytest.mark.parametrize("real,lam_prbfit,lam_laplacian", [(False, 0.0, 0.0), (False, 0.3, 0.05),
                                                           (True, 0.2, 0.05)])
def test_bh_loop_matches_oracle(gpu, real, lam_prbfit, lam_laplacian):
    print(f"[case] BH loop real={real} lam_prbfit={lam_prbfit} lam_laplacian={lam_laplacian}", flush=True)
    rec, orc, start = make_bh(real, lam_prbfit, lam_laplacian)
    rec.BH()
    torch.cuda.synchronize()
    vo, hist = orc.BH(start, rec.niter)
    alphas = [float(a) for a in rec.alphas]
    betas = [float(b) for b in rec.betas]
    errs = [e for i, e, _ in rec.table if i >= 0]
    print("alpha", alphas, [h[0] for h in hist], flush=True)
    print("beta", betas, [h[1] for h in hist], flush=True)
    print("err", errs, [h[2] for h in hist], flush=True)
    for k in range(rec.niter):
        assert alphas[k] == pytest.approx(hist[k][0], rel=1e-4)
        # beta is a ratio whose numerator nearly cancels when the directions
        # are close to conjugate: compared to 1e-3 relative or 1e-6 absolute
        assert betas[k] == pytest.approx(hist[k][1], rel=1e-3, abs=1e-6)
        assert errs[k] == pytest.approx(hist[k][2], rel=1e-4)
    assert all(b < a for a, b in zip(errs, errs[1:]))
    for k in ("obj", "prb", "pos"):
        e = rel(rec.vars[k], vo[k])
        print(k, e, flush=True)
        assert e < 1e-4
    # proj is carried as R(obj) by linearity
    assert rel(rec.vars["proj"], orc.fwd_tomo(vo["obj"] / orc.norm_const)) < 1e-4


def test_bh_gradient_and_hessian_pieces(gpu):
    """The loop's full gradient (fused data term + RT + regulariser + probe fit)
    and its bilinear form against the oracle at one point."""
    print("[case] BH gradients / hessian pieces", flush=True)
    rec, orc, start = make_bh(False, 0.3, 0.05)
    x = {k: torch.from_numpy(v.astype(np.complex64 if v.dtype == np.complex128 else np.float32)).cuda()
         for k, v in start.items()}
    x["obj"] /= float(rec.norm_const)
    x["proj"] = torch.empty_like(rec.vars["proj"])
    rec.vars["obj"].copy_(x["obj"])
    rec.fwd_tomo(x["obj"], out=x["proj"])
    xo = {k: v.cpu().numpy().astype(np.float64 if v.dtype == torch.float32 else np.complex128) for k, v in x.items()}
    g = {k: torch.zeros_like(v) for k, v in x.items()}
    rec.gradients(x, g)
    go = orc.gradients(xo)
    for k in ("obj", "prb", "pos", "proj"):
        assert rel(g[k], go[k]) < (1e-4 if k == "pos" else 2e-5), k
    rng = np.random.default_rng(3)
    e = {k: (0.1 * (crandn(rng, v.shape) if np.iscomplexobj(v) else rng.standard_normal(v.shape)))
         for k, v in xo.items()}
    e["proj"] = orc.fwd_tomo(e["obj"])
    ed = {k: torch.from_numpy(v.astype(np.complex64 if np.iscomplexobj(v) else np.float32)).cuda()
          for k, v in e.items()}
    assert float(rec.hessian_dev(x, ed, ed)) == pytest.approx(orc.hessian(xo, e, e), rel=1e-4)
    assert rec.min(x["prb"], x["obj"], x["pos"], x["proj"]) == pytest.approx(orc.min(xo), rel=1e-5)


def test_bh_ten_iterations_precision(gpu):
    """Precision study of the loop: ten FP32 iterations (fused gradient and
    bilinear forms, FP64 tables and reductions) against the FP64 restatement.
    The objective after every iteration is compared at 1e-4 relative, and the
    variables after ten iterations at 1e-3 relative L2 (small step-size
    differences accumulate along the iterations)."""
    print("[case] BH loop ten iterations precision", flush=True)
    rec, orc, start = make_bh(False, 0.3, 0.05, niter=10, seed=23)
    rec.BH()
    torch.cuda.synchronize()
    vo, hist = orc.BH(start, rec.niter)
    errs = [e for i, e, _ in rec.table if i >= 0]
    rels = [abs(a - h[2]) / abs(h[2]) for a, h in zip(errs, hist)]
    print("objective rel. errors", ["%.2e" % r for r in rels], flush=True)
    assert max(rels) < 1e-4
    for k in ("obj", "prb", "pos"):
        e = rel(rec.vars[k], vo[k])
        print(k, e, flush=True)
        assert e < 1e-3
