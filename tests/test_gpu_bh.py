"""GPU parity of the fused BH gradient (holotomocupy_amd.bh) against the FP64
oracle cascade and the per-operator cascade.  Tolerances as for the cascade:
relative L2 <= 1e-5 (object, probe), 1e-4 (positions, objective)."""
import numpy as np
import pytest
import torch

from test_gpu_rec import make_problem, dev, lst, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("real", [False, True])
@pytest.mark.parametrize("geom", [(32, 40, 3, 4), (64, 48, 2, 3), (16, 24, 4, 5)])
def test_fused_gradient_vs_oracle(gpu, real, geom):
    n, nobj, ndist, B = geom
    print(f"[case] fused gradient n={n} nobj={nobj} ndist={ndist} B={B} real={real}", flush=True)
    rec, cas, x, v, w, d = make_problem(real, n=n, nobj=nobj, ndist=ndist, B=B, seed=20 + n)
    xd, dd = dev(x), torch.from_numpy(d).cuda()
    grads = {"prb": torch.zeros_like(xd["prb"]), "proj": torch.empty_like(xd["proj"]),
             "pos": torch.empty_like(xd["pos"])}
    obj = torch.zeros(1, dtype=torch.float64, device="cuda")
    rec.gradients_fused(xd, grads, dd, objective=obj)
    gprb, gproj, gpos = cas.gradients(*lst(x), d)
    assert rel(grads["prb"], gprb) < 1e-5
    assert rel(grads["proj"], gproj) < 1e-5
    assert rel(grads["pos"], gpos) < 1e-4
    assert obj.item() / rec.data_size == pytest.approx(cas.objective(*lst(x), d), rel=1e-4)
    # the per-operator cascade gives the same gradient
    g2 = {"prb": torch.zeros_like(xd["prb"]), "proj": torch.empty_like(xd["proj"]),
          "pos": torch.empty_like(xd["pos"])}
    rec.gradients_cascade(xd, g2, dd)
    assert rel(grads["proj"], g2["proj"].cpu().numpy()) < 1e-5
    assert rel(grads["prb"], g2["prb"].cpu().numpy()) < 1e-5


def test_fused_gradient_deterministic_and_accumulates(gpu):
    print("[case] fused gradient determinism", flush=True)
    rec, cas, x, v, w, d = make_problem(False, seed=31)
    xd, dd = dev(x), torch.from_numpy(d).cuda()
    outs = []
    for _ in range(2):
        g = {"prb": torch.zeros_like(xd["prb"]), "proj": torch.empty_like(xd["proj"]),
             "pos": torch.empty_like(xd["pos"])}
        rec.gradients_fused(xd, g, dd)
        outs.append(g)
    for k in outs[0]:
        assert torch.equal(outs[0][k], outs[1][k]), k
    # grads['prb'] is accumulated (the reference's +=, rec_mpi.py:347)
    g = {"prb": outs[0]["prb"].clone(), "proj": torch.empty_like(xd["proj"]), "pos": torch.empty_like(xd["pos"])}
    rec.gradients_fused(xd, g, dd)
    assert torch.allclose(g["prb"], 2 * outs[0]["prb"], rtol=1e-6, atol=1e-12)


def test_fused_gradient_two_probes(gpu):
    """Angles of two requests with different probes in one launch: each probe
    gradient equals the one computed separately."""
    print("[case] fused gradient two probes", flush=True)
    from holotomocupy_amd.bh import BHGradient
    rec, cas, x, v, w, d = make_problem(False, seed=41, B=4)
    bh = BHGradient(rec)
    xd, dd = dev(x), torch.from_numpy(d).cuda()
    B = xd["proj"].shape[0]
    prb2 = torch.stack([xd["prb"], xd["prb"] * (0.8 + 0.1j)])        # [2, ndist, nz, n]
    pidx = torch.tensor([0, 0, 1, 1], dtype=torch.int32, device="cuda")
    mag = rec.eff_demagnifications[:B].contiguous()
    gproj = torch.empty_like(xd["proj"])
    gpos = torch.empty_like(xd["pos"])
    gprb = torch.zeros_like(prb2)
    s, rp = bh.scales(B)
    bh(prb2, xd["proj"], xd["pos"], dd, mag, gproj, gpos, gprb, accumulate_prb=False,
       probe_index=pidx, scales=s, rho_prb=rp.repeat(2))
    for p, sl in ((0, slice(0, 2)), (1, slice(2, 4))):
        gp = torch.empty_like(xd["proj"][sl])
        gq = torch.empty_like(xd["pos"][sl])
        gr = torch.zeros_like(xd["prb"])
        bh(prb2[p].contiguous(), xd["proj"][sl].contiguous(), xd["pos"][sl].contiguous(), dd[sl].contiguous(),
           mag[sl].contiguous(), gp, gq, gr, accumulate_prb=False, scales=s[sl].contiguous(), rho_prb=rp)
        assert torch.allclose(gprb[p], gr, rtol=1e-5, atol=1e-7 * gr.abs().max().item())
        assert torch.allclose(gproj[sl], gp, rtol=1e-5, atol=1e-7 * gp.abs().max().item())
        assert torch.allclose(gpos[sl], gq, rtol=1e-5, atol=1e-7 * gq.abs().max().item())


def test_fused_gradient_address_table_equals_packed(gpu):
    """htamd_bh_gradient_tab (each angle's projection, data and object
    gradient and each probe slot's probe and probe gradient addressed through
    a table, as the serving loop calls it) equals htamd_bh_gradient on the
    same inputs packed, bit for bit: two requests of 3 and 1 angles with
    different probes, their arrays in separate allocations."""
    print("[case] fused gradient address table", flush=True)
    from holotomocupy_amd.bh import BHGradient
    rec, cas, x, v, w, d = make_problem(False, seed=43, B=4)
    bh = BHGradient(rec)
    xd, dd = dev(x), torch.from_numpy(d).cuda()
    B = 4
    prb2 = torch.stack([xd["prb"], xd["prb"] * (0.8 + 0.1j)])
    pidx = torch.tensor([0, 0, 0, 1], dtype=torch.int32, device="cuda")
    mag = rec.eff_demagnifications[:B].contiguous()
    s, rp = bh.scales(B)
    rho2 = rp.repeat(2)
    gproj, gpos, gprb = torch.empty_like(xd["proj"]), torch.empty_like(xd["pos"]), torch.empty_like(prb2)
    bh(prb2, xd["proj"], xd["pos"], dd, mag, gproj, gpos, gprb, accumulate_prb=False, probe_index=pidx, scales=s,
       rho_prb=rho2)
    # per-request copies in their own allocations
    reqs = [(slice(0, 3), 0), (slice(3, 4), 1)]
    proj_r = [xd["proj"][sl].clone() for sl, _ in reqs]
    data_r = [dd[sl].clone() for sl, _ in reqs]
    prb_r = [prb2[p].clone() for _, p in reqs]
    gproj_r = [torch.full_like(t, float("nan")) for t in proj_r]
    gprb_r = [torch.full_like(t, float("nan")) for t in prb_r]
    cols = [[], [], [], [], []]
    for k in range(2):
        for t, c in ((proj_r[k], 0), (data_r[k], 1), (gproj_r[k], 2)):
            step = t[0].numel() * t.element_size()
            cols[c].extend(t.data_ptr() + i * step for i in range(t.shape[0]))
        cols[3].append(prb_r[k].data_ptr())
        cols[4].append(gprb_r[k].data_ptr())
    tab = torch.tensor(sum(cols, []), dtype=torch.int64, device="cuda")
    gpos_t = torch.empty_like(gpos)
    ws = torch.empty(bh.workspace_bytes(B), dtype=torch.uint8, device="cuda")
    bh.call_tab(tab, B, 2, xd["pos"], mag, pidx, s, rho2, gpos_t, ws)
    torch.cuda.synchronize()
    assert torch.equal(gpos_t, gpos)
    for k, (sl, p) in enumerate(reqs):
        assert torch.equal(gproj_r[k], gproj[sl]), k
        assert torch.equal(gprb_r[k], gprb[p]), k


def test_fused_gradient_nonsquare_detector(gpu):
    """nz != n and magnified shifts."""
    print("[case] fused gradient non-square", flush=True)
    from types import SimpleNamespace
    from holotomocupy_amd.rec import Rec, physics
    from oracle import cascade as oc
    from oracle import propagation as op
    from oracle import shift as osh
    rng = np.random.default_rng(5)
    n, nz, nobj, nzobj, ndist, B = 64, 32, 72, 40, 2, 3
    args = SimpleNamespace(energy=17.1, detector_pixelsize=1.4760147601476e-6 * 16, focustodetectordistance=1.217,
                           z1=np.array([5.110, 6.879]) * 1e-3, ndist=ndist, ntheta=B, nz=nz, n=n, nzobj=nzobj,
                           nobj=nobj, obj_dtype="complex64", rho=[1, 0.5, 0.2])
    rec = Rec(args)
    ph = physics(args.energy, args.z1, args.focustodetectordistance, args.detector_pixelsize)
    cas = oc.Cascade(op.Propagation(n, nz, B, ndist, ph.wavelength, ph.voxelsize, ph.distance),
                     osh.Shift(n, nobj, nz, nzobj), ndist, nz, n, nzobj, nobj, rec.data_size,
                     np.tile(1 / ph.norm_magnifications, (B, 1)), rho=args.rho)
    prb = (1 + 0.1 * (rng.standard_normal((ndist, nz, n)) + 1j * rng.standard_normal((ndist, nz, n)))).astype(np.complex64)
    proj = (0.3 * (rng.standard_normal((B, nzobj, nobj)) + 1j * rng.standard_normal((B, nzobj, nobj)))).astype(np.complex64)
    pos = rng.uniform(-3, 3, (B, ndist, 2)).astype(np.float32)
    d = (np.abs(cas.forward(prb.astype(np.complex128), proj.astype(np.complex128), pos.astype(np.float64))) * 1.1).astype(np.float32)
    x = {"prb": prb, "proj": proj, "pos": pos}
    xd, dd = dev(x), torch.from_numpy(d).cuda()
    g = {"prb": torch.zeros_like(xd["prb"]), "proj": torch.empty_like(xd["proj"]), "pos": torch.empty_like(xd["pos"])}
    rec.gradients_fused(xd, g, dd)
    gprb, gproj, gpos = cas.gradients(*lst(x), d)
    assert rel(g["prb"], gprb) < 1e-5
    assert rel(g["proj"], gproj) < 1e-5
    assert rel(g["pos"], gpos) < 1e-4


@pytest.mark.parametrize("real", [False, True])
def test_fused_gradient_magnification_below_one(gpu, real):
    """A negative shrink puts some magnifications below 1 (wider candidate
    brackets in the gather-form Sadj); the rest stay at or above 1."""
    print(f"[case] fused gradient with magnifications below one real={real}", flush=True)
    B, nd = 3, 3
    shrink = np.array([[-0.3, 0.0, 0.1], [-0.15, -0.4, 0.0], [0.2, -0.05, -0.35]], dtype=np.float64)
    rec, cas, x, v, w, d = make_problem(real, n=32, nobj=40, ndist=nd, B=B, seed=31, shrink=shrink)
    assert float(rec.eff_demagnifications.min()) < 1.0 <= float(rec.eff_demagnifications.max())
    xd, dd = dev(x), torch.from_numpy(d).cuda()
    grads = {"prb": torch.zeros_like(xd["prb"]), "proj": torch.empty_like(xd["proj"]),
             "pos": torch.empty_like(xd["pos"])}
    rec.gradients_fused(xd, grads, dd)
    gprb, gproj, gpos = cas.gradients(*lst(x), d)
    assert rel(grads["prb"], gprb) < 1e-5
    assert rel(grads["proj"], gproj) < 1e-5
    assert rel(grads["pos"], gpos) < 1e-4


def test_fused_gradient_bench_size_properties(gpu):
    """Detector 2048^2, object plane 1024^2, 4 distances, 4 angles (the bench's
    shapes): the fused gradient agrees with the per-operator cascade (a second,
    independent implementation) to the operator bar, is deterministic, and its
    per-angle outputs do not depend on how the angles are split across
    launches (the probe gradient, a sum over angles, to 1e-6)."""
    print("[case] fused gradient 2048^2 properties", flush=True)
    from types import SimpleNamespace
    from holotomocupy_amd import phantom
    from holotomocupy_amd.rec import Rec
    n, nobj, nd, B = 2048, 1024, 4, 4
    rng = np.random.default_rng(19)
    args = SimpleNamespace(energy=17.1, detector_pixelsize=1.4760147601476e-6 * 2, focustodetectordistance=1.217,
                           z1=np.array([5.110, 5.464, 6.879, 9.817]) * 1e-3, ndist=nd, ntheta=B, nz=n, n=n,
                           nzobj=nobj, nobj=nobj, obj_dtype="complex64", rho=[1.0, 0.05, 0.02])
    rec = Rec(args)
    x = {"prb": torch.from_numpy(phantom.probe(rng, nd, n, n)).cuda(),
         "proj": torch.from_numpy(phantom.projections(rng, B, nobj, nobj)).cuda(),
         "pos": torch.from_numpy(phantom.positions(rng, B, nd)).cuda()}
    with torch.no_grad():
        data = (rec.forward(x).abs() * (1 + 0.05 * torch.randn((B, nd, n, n), device="cuda"))).float().contiguous()

    def zeros():
        return {"prb": torch.zeros_like(x["prb"]), "proj": torch.empty_like(x["proj"]),
                "pos": torch.empty_like(x["pos"])}

    g1, g2, gc = zeros(), zeros(), zeros()
    rec.gradients_fused(x, g1, data)
    rec.gradients_fused(x, g2, data)
    for k in g1:
        assert torch.equal(g1[k], g2[k]), k
    rec.gradients_cascade(x, gc, data)
    print({k: rel(g1[k], gc[k].cpu().numpy()) for k in g1}, flush=True)
    assert rel(g1["prb"], gc["prb"].cpu().numpy()) < 1e-5
    assert rel(g1["proj"], gc["proj"].cpu().numpy()) < 1e-5
    assert rel(g1["pos"], gc["pos"].cpu().numpy()) < 1e-4
    gs = zeros()
    for st in (0, 2):
        part = {"prb": x["prb"], "proj": x["proj"][st:st + 2], "pos": x["pos"][st:st + 2]}
        out = {"prb": gs["prb"], "proj": gs["proj"][st:st + 2], "pos": gs["pos"][st:st + 2]}
        rec.gradients_fused(part, out, data[st:st + 2], st=st)
    assert torch.equal(gs["proj"], g1["proj"]) and torch.equal(gs["pos"], g1["pos"])
    assert rel(gs["prb"], g1["prb"].cpu().numpy()) < 1e-6
