"""torch.autograd bindings of the hand-written adjoints.

* ``propagate(psi, j, prop)``: D_j with backward DT_j (the propagator is
  linear; its adjoint is the conjugate-kernel operator, propagation.py:91-112).
* ``radon(obj, tomo)``: R with backward RT (tomo.py:56-116).
* ``shift(c, r, m, sh)``: S(c, r, m) with backward Sadj for c (shift.py:102-122);
  r and m are treated as constants.
* ``bh_objective(prb, proj, pos, data, rec, st=0)``: the data term
  sum (|D(prb exp(i S(coeff proj)))| - d)^2 / data_size of Rec (rec_mpi.py:517-527,
  815-835) whose backward is the fused BH gradient kernel (htamd_bh_gradient,
  rec_mpi.py:327-353 without the rho^2 scaling).
* ``bh_bilinear(prb, proj, pos, data, rec, dir1, dir2=None)``: the BH bilinear
  form of that objective (its second derivative along two directions) through
  the fused bilinear-form kernel; the BH step sizes are ratios of these.

For a real-valued loss of complex inputs PyTorch's gradient is
dL/dx + i dL/dy, i.e. the g with dL = Re<g, dz> -- the convention of the
reference's gradients -- so the kernels' outputs are passed through unchanged.
"""
import torch

from . import _lib


class _Propagate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, psi, j, prop):
        ctx.j, ctx.prop = j, prop
        return prop.D(psi, j)

    @staticmethod
    def backward(ctx, g):
        return ctx.prop.DT(g.contiguous(), ctx.j), None, None


def propagate(psi, j, prop):
    return _Propagate.apply(psi, j, prop)


class _Radon(torch.autograd.Function):
    @staticmethod
    def forward(ctx, obj, tomo):
        ctx.tomo = tomo
        return tomo.R(obj)

    @staticmethod
    def backward(ctx, g):
        return ctx.tomo.RT(g.contiguous()), None


def radon(obj, tomo):
    return _Radon.apply(obj, tomo)


class _Shift(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c, r, m, sh):
        ctx.save_for_backward(r, m)
        ctx.sh = sh
        return sh.curlySc(c, r, m)

    @staticmethod
    def backward(ctx, g):
        r, m = ctx.saved_tensors
        return ctx.sh.Sadj(g.contiguous(), r, m), None, None, None


def shift(c, r, m, sh):
    return _Shift.apply(c, r, m, sh)


class _BHObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, prb, proj, pos, data, rec, st):
        from .bh import BHGradient
        bh = getattr(rec, "_bh", None)
        if bh is None:
            bh = rec._bh = BHGradient(rec)
        B = proj.shape[0]
        prb, proj, pos, data = (t.contiguous() for t in (prb, proj, pos, data))
        mag = rec.eff_demagnifications[st:st + B].contiguous()
        gprb = torch.zeros_like(prb)
        gproj = torch.empty_like(proj)
        gpos = torch.empty_like(pos)
        obj = torch.zeros(1, dtype=torch.float64, device=prb.device)
        # unit rho: the plain gradient of the data term
        scales = torch.tensor([2.0 / rec.data_size, 1.0, 1.0, 0.0], dtype=torch.float32,
                              device=prb.device).repeat(B, 1)
        rho_prb = torch.ones(1, dtype=torch.float32, device=prb.device)
        bh(prb, proj, pos, data, mag, gproj, gpos, gprb, accumulate_prb=False, objective=obj, scales=scales,
           rho_prb=rho_prb)
        ctx.save_for_backward(gprb, gproj, gpos)
        return obj[0] / rec.data_size

    @staticmethod
    def backward(ctx, gout):
        gprb, gproj, gpos = ctx.saved_tensors
        s = gout.to(torch.float32)
        return gprb * s, gproj * s, gpos * s, None, None, None


def bh_objective(prb, proj, pos, data, rec, st=0):
    """Data term of the BH objective over the angles [st, st + B) with the
    fused gradient as its backward.  The gradient is computed together with
    the value (one fused launch sequence) and scaled by the incoming
    gradient in backward."""
    _lib.require_gpu()
    return _BHObjective.apply(prb, proj, pos, data, rec, st)


def bh_bilinear(prb, proj, pos, data, rec, dir1, dir2=None, st=0):
    """Second derivative of the data term that bh_objective returns,
    d^2 f(x)(y, z), through the fused bilinear-form kernel (htamd_bh_hessian,
    Rec.hessian_cascade of rec_mpi.py:250-303): x = (prb, proj, pos), y = dir1
    and z = dir2 (None: z = y), each a (dprb, dproj, dpos) triple.  The
    reference's d2s pairing is kept, so for y != z this is the reference's
    form, exact for y = z.  Returns a float64 device tensor [1]."""
    _lib.require_gpu()
    c = lambda t: t.detach().contiguous()  # noqa: E731
    x = {"prb": c(prb), "proj": c(proj), "pos": c(pos)}
    y = {"prb": c(dir1[0]), "proj": c(dir1[1]), "pos": c(dir1[2])}
    z = y if dir2 is None else {"prb": c(dir2[0]), "proj": c(dir2[1]), "pos": c(dir2[2])}
    return rec.hessian_fused(x, y, z, c(data), st)
