"""GPU parity of the fused BH bilinear form (htamd_bh_hessian, Rec.hessian_fused)
against the FP64 oracle cascade (oracle/cascade.py, pinned by second
differences and symmetry in tests/test_oracle_cascade.py).

Tolerance: 1e-4 relative for the scalar (a sum of many signed FP32 terms,
accumulated in FP64), as for the per-operator bilinear form.  At the bench
size the checks are size-independent: the equal-direction form against the
general form with a copy of the direction, symmetry in the two directions,
agreement with the per-operator cascade, and determinism."""
import numpy as np
import pytest
import torch

from test_gpu_rec import dev, lst, make_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("real", [False, True])
@pytest.mark.parametrize("shape", [(32, 40, 3, 4), (64, 32, 2, 3)])
def test_fused_hessian_matches_oracle(gpu, real, shape):
    n, nobj, ndist, B = shape
    print(f"[case] fused hessian real={real} n={n} nobj={nobj} ndist={ndist} B={B}", flush=True)
    rec, cas, x, v, w, d = make_problem(real, n=n, nobj=nobj, ndist=ndist, B=B, seed=7)
    xd, vd, wd, dd = dev(x), dev(v), dev(w), torch.from_numpy(d).cuda()
    h_vv = float(rec.hessian_fused(xd, vd, vd, dd))
    h_vw = float(rec.hessian_fused(xd, vd, wd, dd))
    lx, lv, lw = lst(x), lst(v), lst(w)
    ref_vv, ref_vw = cas.hessian(lx, lv, lv, d), cas.hessian(lx, lv, lw, d)
    print("vv", h_vv, ref_vv, "vw", h_vw, ref_vw, flush=True)
    assert h_vv == pytest.approx(ref_vv, rel=1e-4)
    assert h_vw == pytest.approx(ref_vw, rel=1e-4)


def test_fused_hessian_accumulates_and_is_deterministic(gpu):
    print("[case] fused hessian accumulate / determinism", flush=True)
    rec, cas, x, v, w, d = make_problem(False, seed=8)
    xd, vd, wd, dd = dev(x), dev(v), dev(w), torch.from_numpy(d).cuda()
    a = rec.hessian_fused(xd, vd, wd, dd)
    b = rec.hessian_fused(xd, vd, wd, dd)
    assert float(a) == float(b)
    acc = torch.tensor([1.5], dtype=torch.float64, device="cuda")
    rec.hessian_fused(xd, vd, wd, dd, out=acc)
    assert float(acc) == pytest.approx(1.5 + float(a), rel=1e-12)


def test_fused_hessian_bench_size_properties(gpu):
    """2048^2 detector, 1024^2 object plane, 4 distances, 2 angles."""
    print("[case] fused hessian 2048^2 properties", flush=True)
    from types import SimpleNamespace
    from holotomocupy_amd import phantom
    from holotomocupy_amd.rec import Rec
    n, nobj, nd, B = 2048, 1024, 4, 2
    rng = np.random.default_rng(9)
    args = SimpleNamespace(energy=17.1, detector_pixelsize=1.4760147601476e-6 * 2, focustodetectordistance=1.217,
                           z1=np.array([5.110, 5.464, 6.879, 9.817]) * 1e-3, ndist=nd, ntheta=B, nz=n, n=n,
                           nzobj=nobj, nobj=nobj, obj_dtype="complex64", rho=[1.0, 0.05, 0.02])
    rec = Rec(args)
    x = {"prb": torch.from_numpy(phantom.probe(rng, nd, n, n)).cuda(),
         "proj": torch.from_numpy(phantom.projections(rng, B, nobj, nobj)).cuda(),
         "pos": torch.from_numpy(phantom.positions(rng, B, nd)).cuda()}
    with torch.no_grad():
        data = (rec.forward(x).abs() * 1.01).float().contiguous()
    g = torch.Generator(device="cuda").manual_seed(9)

    def direction():
        return {"prb": 0.05 * torch.randn(x["prb"].shape, dtype=torch.complex64, device="cuda", generator=g),
                "proj": 0.05 * torch.randn(x["proj"].shape, dtype=torch.complex64, device="cuda", generator=g),
                "pos": 0.3 * torch.randn(x["pos"].shape, dtype=torch.float32, device="cuda", generator=g)}

    v, w = direction(), direction()
    vc = {k: t.clone() for k, t in v.items()}
    h_same = float(rec.hessian_fused(x, v, v, data))
    h_copy = float(rec.hessian_fused(x, v, vc, data))
    assert h_copy == pytest.approx(h_same, rel=1e-5)
    h_vw, h_wv = float(rec.hessian_fused(x, v, w, data)), float(rec.hessian_fused(x, w, v, data))
    assert h_vw == pytest.approx(h_wv, rel=1e-5)
    h_plain = rec.hessian_cascade(x, v, w, data)
    print("same", h_same, "copy", h_copy, "vw", h_vw, "wv", h_wv, "plain", h_plain, flush=True)
    assert h_vw == pytest.approx(h_plain, rel=1e-4)


@pytest.mark.parametrize("real", [False, True])
def test_pair_form_matches_the_two_forms(gpu, real):
    """htamd_bh_hessian_pair: (H(v, w), H(w, w)) in one pass equals the two
    separate fused forms (same arithmetic, regrouped: 1e-5) and the oracle."""
    print(f"[case] fused hessian pair real={real}", flush=True)
    rec, cas, x, v, w, d = make_problem(real, seed=10)
    xd, vd, wd, dd = dev(x), dev(v), dev(w), torch.from_numpy(d).cuda()
    pair = rec.hessian_pair_fused(xd, vd, wd, dd)
    h_vw = float(rec.hessian_fused(xd, vd, wd, dd))
    h_ww = float(rec.hessian_fused(xd, wd, wd, dd))
    assert float(pair[0]) == pytest.approx(h_vw, rel=1e-5)
    assert float(pair[1]) == pytest.approx(h_ww, rel=1e-5)
    lx, lv, lw = lst(x), lst(v), lst(w)
    assert float(pair[0]) == pytest.approx(cas.hessian(lx, lv, lw, d), rel=1e-4)
    assert float(pair[1]) == pytest.approx(cas.hessian(lx, lw, lw, d), rel=1e-4)
