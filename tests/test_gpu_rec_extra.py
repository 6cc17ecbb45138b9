"""Probe-fit term, chunked tomography (fwd_tomo / gF4) of Rec on the GPU.

Pins: the probe-fit gradient against central differences of the FP64 oracle
objective lam/prb_size || |D prb| - ref ||^2, its bilinear form against
second differences; fwd_tomo / gF4 over host arrays (chunked through
Chunking.gpu_batch) against the FP64 oracle Radon transform at 1e-5."""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from oracle import propagation as op
from oracle import tomo as ot

pytestmark = pytest.mark.gpu


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def make(n=32, nobj=32, ndist=2, ntheta=6, nchunk=3, seed=11, lam=0.7):
    from holotomocupy_amd.rec import Rec, physics
    rng = np.random.default_rng(seed)
    theta = (np.arange(ntheta) * np.pi / ntheta).astype(np.float32).astype(np.float64)
    args = SimpleNamespace(energy=17.1, detector_pixelsize=1.4760147601476e-6 * 16,
                           focustodetectordistance=1.217, z1=np.array([5.110, 5.464])[:ndist] * 1e-3,
                           ndist=ndist, ntheta=ntheta, nz=n, n=n, nzobj=nobj, nobj=nobj, nchunk=nchunk,
                           obj_dtype="complex64", rho=[1, 0.5, 0.2], lam_prbfit=lam, theta=theta, mask_r=0)
    ph = physics(args.energy, args.z1, args.focustodetectordistance, args.detector_pixelsize)
    prop = op.Propagation(n, n, 1, ndist, ph.wavelength, ph.voxelsize, ph.distance)
    yy, xx = np.mgrid[0:n, 0:n] - n / 2
    prb = np.stack([np.exp(-(xx ** 2 + yy ** 2) / (2 * (0.4 * n) ** 2)) * np.exp(0.3j * np.cos(0.2 * xx + j))
                    for j in range(ndist)]).astype(np.complex64)
    ref = np.stack([np.abs(prop.D(prb[j:j + 1].astype(np.complex128), j))[0] for j in range(ndist)])
    ref = (ref * (1 + 0.05 * rng.standard_normal(ref.shape))).astype(np.float32)
    args.ref = ref
    rec = Rec(args)
    return rec, prop, prb, ref, rng, theta


def prbfit_objective(prop, prb, ref, lam):
    s = 0.0
    for j in range(prb.shape[0]):
        s += np.sum((np.abs(prop.D(prb[j:j + 1], j)) - ref[j:j + 1]) ** 2)
    return lam * s / prb.size


def test_prbfit_gradient_and_hessian(gpu):
    print("[case] probe-fit gradient / hessian", flush=True)
    rec, prop, prb, ref, rng, _ = make()
    v = (0.1 * crandn(rng, prb.shape)).astype(np.complex64)
    g = torch.zeros_like(torch.from_numpy(prb)).cuda()
    rec.gradient_prbfit(g, torch.from_numpy(prb).cuda())
    g = g.cpu().numpy().astype(np.complex128) / rec.rho_sq["prb"]
    p64, v64, r64 = prb.astype(np.complex128), v.astype(np.complex128), ref.astype(np.float64)
    h = 1e-3
    fd = (prbfit_objective(prop, p64 + h * v64, r64, rec.lam_prbfit)
          - prbfit_objective(prop, p64 - h * v64, r64, rec.lam_prbfit)) / (2 * h)
    an = np.sum(g.real * v64.real + g.imag * v64.imag)
    assert abs(an - fd) < 1e-4 * abs(fd)
    hs = rec.hessian_prbfit(torch.from_numpy(prb).cuda(), torch.from_numpy(v).cuda(), torch.from_numpy(v).cuda())
    h2 = 1e-2
    sd = (prbfit_objective(prop, p64 + h2 * v64, r64, rec.lam_prbfit) - 2 * prbfit_objective(prop, p64, r64, rec.lam_prbfit)
          + prbfit_objective(prop, p64 - h2 * v64, r64, rec.lam_prbfit)) / h2 ** 2
    assert abs(hs - sd) < 1e-3 * abs(sd)


def test_fwd_tomo_and_gF4_chunked_host_arrays(gpu):
    print("[case] fwd_tomo / gF4 through gpu_batch", flush=True)
    rec, _, _, _, rng, theta = make(nobj=32, ntheta=6, nchunk=3)
    nobj = 32
    u = crandn(rng, (nobj, nobj, nobj)).astype(np.complex64)  # [nzobj, nobj, nobj]
    out = np.zeros((len(theta), nobj, nobj), dtype=np.complex64)
    rec.fwd_tomo(u, out)
    o = ot.Tomo(nobj, nobj, theta, 0)
    ref = o.R(u.astype(np.complex128))
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-5
    gp = crandn(rng, (len(theta), nobj, nobj)).astype(np.complex64)
    gu = np.zeros((nobj, nobj, nobj), dtype=np.complex64)
    rec.gF4(gu, gp)
    ref2 = o.RT(gp.astype(np.complex128))
    assert np.linalg.norm(gu - ref2) / np.linalg.norm(ref2) < 1e-5


def test_fwd_tomo_and_gF4_device_slabs(gpu, monkeypatch):
    """Device-resident fwd_tomo / gF4 in slabs that do not divide the slice
    count (8 + 8 + 4 of 20 slices): equal bit for bit to one whole-volume
    slab, and to the FP64 oracle at 1e-5; a float64 volume goes through the
    slab loop (cast to float32 per slab) instead of the direct-output path."""
    from holotomocupy_amd.rec import Rec
    print("[case] fwd_tomo / gF4 device slabs 8 of 20", flush=True)
    rec, _, _, _, rng, theta = make(nobj=32, ntheta=6, nchunk=3)
    rec.nzobj = nzobj = 20
    nobj, nth = 32, len(theta)
    u = torch.from_numpy(crandn(rng, (nzobj, nobj, nobj)).astype(np.complex64)).cuda()
    gp = torch.from_numpy(crandn(rng, (nth, nzobj, nobj)).astype(np.complex64)).cuda()
    res = {}
    for slab in (nzobj, 8):
        monkeypatch.setattr(Rec, "_tomo_slab", lambda self, s=slab: s)
        out = torch.empty((nth, nzobj, nobj), dtype=torch.complex64, device="cuda")
        rec.fwd_tomo(u, out)
        gu = torch.empty((nzobj, nobj, nobj), dtype=torch.complex64, device="cuda")
        rec.gF4(gu, gp)
        torch.cuda.synchronize()
        res[slab] = (out.clone(), gu.clone())
    assert torch.equal(res[8][0], res[nzobj][0]) and torch.equal(res[8][1], res[nzobj][1])
    o = ot.Tomo(nobj, nzobj, theta, 0)
    ref_r = o.R(u.cpu().numpy().astype(np.complex128))
    ref_rt = o.RT(gp.cpu().numpy().astype(np.complex128))
    assert np.linalg.norm(res[8][0].cpu().numpy() - ref_r) / np.linalg.norm(ref_r) < 1e-5
    assert np.linalg.norm(res[8][1].cpu().numpy() - ref_rt) / np.linalg.norm(ref_rt) < 1e-5
    # a float64 real volume into a float64 output: the per-slab loop casts each slab
    ur = torch.from_numpy(rng.standard_normal((nzobj, nobj, nobj))).cuda()
    outr = torch.empty((nth, nzobj, nobj), dtype=torch.float64, device="cuda")
    rec.fwd_tomo(ur, outr)
    ref_rr = o.R(ur.cpu().numpy())
    assert np.linalg.norm(outr.cpu().numpy() - ref_rr) / np.linalg.norm(ref_rr) < 1e-5
