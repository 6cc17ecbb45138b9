"""torch.autograd bindings: the BH objective's backward is the fused gradient
(checked against the FP64 oracle cascade gradient, rho divided out, at 1e-5
like the fused-gradient tests), and the linear operators' backward passes are
their adjoints (checked against the direct adjoint calls)."""
import numpy as np
import pytest
import torch

from test_gpu_rec import make_problem, dev, lst, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("real", [False, True])
def test_bh_objective_backward_is_fused_gradient(gpu, real):
    from holotomocupy_amd.autograd import bh_objective
    print(f"[case] autograd bh_objective real={real}", flush=True)
    rec, cas, x, v, w, d = make_problem(real, seed=61)
    xd = {k: t.clone().requires_grad_(True) for k, t in dev(x).items()}
    dd = torch.from_numpy(d).cuda()
    L = bh_objective(xd["prb"], xd["proj"], xd["pos"], dd, rec)
    assert L.item() == pytest.approx(cas.objective(*lst(x), d), rel=1e-4)
    gprb, gproj, gpos = torch.autograd.grad(3.0 * L, [xd["prb"], xd["proj"], xd["pos"]])
    rp, rpr, rps = cas.gradients(*lst(x), d)
    rho = rec.rho_sq
    assert rel(gprb / 3.0, rp / rho["prb"]) < 1e-5
    assert rel(gproj / 3.0, rpr / rho["obj"]) < 1e-5
    assert rel(gpos / 3.0, rps / rho["pos"]) < 1e-4


def test_linear_operator_backward_is_adjoint(gpu):
    from holotomocupy_amd import Propagation, Shift, Tomo
    from holotomocupy_amd.autograd import propagate, radon, shift
    print("[case] autograd D / R / S backward", flush=True)
    g = torch.Generator(device="cuda").manual_seed(3)
    prop = Propagation(32, 32, 2, 2, 1e-10, 1e-7, [1e-3, 2e-3])
    x = torch.randn((2, 32, 32), dtype=torch.complex64, device="cuda", generator=g).requires_grad_(True)
    y = torch.randn((2, 32, 32), dtype=torch.complex64, device="cuda", generator=g)
    L = torch.sum(torch.real(propagate(x, 1, prop) * torch.conj(y)))
    (gx,) = torch.autograd.grad(L, [x])
    assert torch.allclose(gx, prop.DT(y, 1), rtol=1e-6, atol=1e-6)
    th = np.arange(12) * np.pi / 12
    tomo = Tomo(32, 2, th, 0)
    u = torch.randn((2, 32, 32), dtype=torch.complex64, device="cuda", generator=g).requires_grad_(True)
    s = torch.randn((12, 2, 32), dtype=torch.complex64, device="cuda", generator=g)
    L = torch.sum(torch.real(radon(u, tomo) * torch.conj(s)))
    (gu,) = torch.autograd.grad(L, [u])
    assert torch.allclose(gu, tomo.RT(s), rtol=1e-6, atol=1e-6)
    sh = Shift(32, 40, 32, 40, "complex64")
    c = torch.randn((2, 40, 40), dtype=torch.complex64, device="cuda", generator=g).requires_grad_(True)
    r = torch.tensor([[0.3, -1.2], [2.5, 0.7]], dtype=torch.float32, device="cuda")
    m = torch.tensor([0.9, 1.1], dtype=torch.float32, device="cuda")
    q = torch.randn((2, 32, 32), dtype=torch.complex64, device="cuda", generator=g)
    L = torch.sum(torch.real(shift(c, r, m, sh) * torch.conj(q)))
    (gc,) = torch.autograd.grad(L, [c])
    assert torch.allclose(gc, sh.Sadj(q, r, m), rtol=1e-6, atol=1e-6)


def test_bh_bilinear_is_derivative_of_autograd_gradient(gpu):
    """bh_bilinear(x; v, v) against central differences (with Richardson
    extrapolation) of <grad bh_objective(x + t v), v> at t = 0, the gradient
    coming from autograd; and against the oracle's bilinear form for v != w.
    The differences of FP32 gradients at step 0.05 carry about 2e-3 of
    rounding (the FP64 oracle's own truncation error there is 6e-6), so that
    check is at 5e-3; the oracle check is the 1e-4 one."""
    from holotomocupy_amd.autograd import bh_bilinear, bh_objective
    print("[case] autograd bh_bilinear", flush=True)
    rec, cas, x, v, w, d = make_problem(False, seed=62)
    xd, vd, wd = dev(x), dev(v), dev(w)
    dd = torch.from_numpy(d).cuda()
    keys = ("prb", "proj", "pos")

    def dgrad(t):
        xs = [(xd[k] + t * vd[k]).requires_grad_(True) for k in keys]
        gs = torch.autograd.grad(bh_objective(*xs, dd, rec), xs)
        tot = 0.0
        for g, k in zip(gs, keys):
            if g.is_complex():
                tot += float(torch.sum(torch.real(torch.conj(g.to(torch.complex128)) * vd[k].to(torch.complex128))))
            else:
                tot += float(torch.sum(g.double() * vd[k].double()))
        return tot

    def central(h):
        return (dgrad(h) - dgrad(-h)) / (2 * h)

    h = 0.05
    fd = (4 * central(h / 2) - central(h)) / 3
    hvv = float(bh_bilinear(xd["prb"], xd["proj"], xd["pos"], dd, rec, (vd["prb"], vd["proj"], vd["pos"])))
    print("bilinear", hvv, "central differences", fd, flush=True)
    assert hvv == pytest.approx(fd, rel=5e-3)
    hvw = float(bh_bilinear(xd["prb"], xd["proj"], xd["pos"], dd, rec, (vd["prb"], vd["proj"], vd["pos"]),
                            (wd["prb"], wd["proj"], wd["pos"])))
    assert hvw == pytest.approx(cas.hessian(lst(x), lst(v), lst(w), d), rel=1e-4)
