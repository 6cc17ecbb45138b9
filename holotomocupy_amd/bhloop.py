"""The BH reconstruction loop of holotomocupy's Rec, kept on the device.

Drop-in for Rec.BH and what it calls (src/holotomocupy/rec_mpi.py:103-325,
400-507, 815-902) on one GPU: the whole problem lives in HBM (object, the
per-angle projections, the data, gradients and directions), so the reference's
pinned host buffers, MPI redistribution and all-reduces fall away.  The
gradient of each chunk of angles is the fused three-pass HIP pipeline and the
bilinear form the fused two-pass one (fused_hessian = False selects the
per-operator cascade instead); R / RT are the USFFT kernels;
the regulariser is the library's 3-D Laplacian.  The step sizes alpha and beta
stay on the device: every inner product accumulates into float64 device
scalars and the axpy updates read their coefficients there, so an iteration
never waits on the host (the reference synchronises for each .get()).

``BHLoop`` is mixed into ``holotomocupy_amd.rec.Rec``.
"""
import csv
import logging
import os
import time

import numpy as np
import torch

from . import _lib

log = logging.getLogger(__name__)


class BHLoop:
    # ------------------------------------------------------------------ memory
    def alloc_arrays(self):
        """Device-resident variables, gradients, directions, data and probe
        reference (the shapes of rec_mpi.py:103-136)."""
        dev = self.device
        odt = torch.float32 if self.obj_real else torch.complex64
        prb_shape = (self.ndist, self.nz, self.n)

        def group():
            return {"obj": torch.zeros((self.nzobj, self.nobj, self.nobj), dtype=odt, device=dev),
                    "pos": torch.zeros((self.ntheta, self.ndist, 2), dtype=torch.float32, device=dev),
                    "prb": torch.zeros(prb_shape, dtype=torch.complex64, device=dev),
                    "proj": torch.zeros((self.ntheta, self.nzobj, self.nobj), dtype=odt, device=dev)}

        self.vars, self.grads, self.etas = group(), group(), group()
        self.data = torch.zeros((self.ntheta, self.ndist, self.nz, self.n), dtype=torch.float32, device=dev)
        self.ref = torch.zeros(prb_shape, dtype=torch.float32, device=dev)
        self.table = []

    # ------------------------------------------------------------------ helpers
    def _nchunk(self):
        return int(getattr(self, "nchunk", self.ntheta))

    def _chunks(self):
        B = self._nchunk()
        for st in range(0, self.ntheta, B):
            yield st, min(B, self.ntheta - st)

    def _blas_ws(self):
        if getattr(self, "_b1ws", None) is None:
            self._b1ws = torch.empty(1024, dtype=torch.float64, device=self.device)
        return self._b1ws

    def _scalar(self, value=0.0):
        return torch.full((1,), float(value), dtype=torch.float64, device=self.device)

    @staticmethod
    def _floats(t):
        return torch.view_as_real(t) if t.is_complex() else t

    def _redot_dev(self, x, y, out, accumulate=False):
        xf, yf = self._floats(x), self._floats(y)
        ws = self._blas_ws()
        _lib.call("htamd_redot", _lib.ptr(xf), _lib.ptr(yf), xf.numel(), _lib.ptr(out), int(accumulate),
                  _lib.ptr(ws), ws.numel() * 8, _lib.stream_ptr())
        return out

    def _axpby_dc(self, x, y, ab, out):
        """out = a x + b y with (a, b) = ab, two float64 device scalars."""
        _lib.call("htamd_axpby_dc", _lib.ptr(self._floats(x)), _lib.ptr(self._floats(y)), _lib.ptr(ab),
                  _lib.ptr(self._floats(out)), self._floats(x).numel(), _lib.stream_ptr())

    def _axpby_redot_dc(self, x, y, ab, result, scale):
        """x = a x + b y, then result += scale Re<y, x> (rec_mpi.py:198-205)."""
        ws = self._blas_ws()
        _lib.call("htamd_axpby_redot_dc", _lib.ptr(self._floats(x)), _lib.ptr(self._floats(y)), _lib.ptr(ab),
                  self._floats(x).numel(), _lib.ptr(result), float(scale), 1, _lib.ptr(ws), ws.numel() * 8,
                  _lib.stream_ptr())

    def _lap(self, u, z0, nz_out, add=None, scale=1.0, out=None):
        """out[k] = add[k] + scale lap(u)[z0 + k] (zero outside u's slices)."""
        if out is None:
            out = torch.empty((nz_out,) + tuple(u.shape[1:]), dtype=u.dtype, device=self.device)
        _lib.call("htamd_laplacian3d", _lib.ptr(u), u.shape[0], u.shape[1], u.shape[2], 1 if u.is_complex() else 0,
                  int(z0), int(nz_out), _lib.ptr(add) if add is not None else None, float(scale), _lib.ptr(out),
                  _lib.stream_ptr())
        return out

    def _biharmonic(self, u, add=None, scale=1.0, out=None):
        """add + scale lap(lap(u_pad)) on u's slices, with the inner Laplacian
        also on the two ghost slices (rec_mpi.py:432-436)."""
        L = self._lap(u, -1, u.shape[0] + 2)
        return self._lap(L, 1, u.shape[0], add=add, scale=scale, out=out)

    # ------------------------------------------------------------------ regulariser (rec_mpi.py:400-494)
    def gradient_laplacian(self, grad_obj, u=None):
        """grad_obj += 2 lam/obj_size (lap)^2 u  (rec_mpi.py:418-443)."""
        lam = getattr(self, "lam_laplacian", 0)
        if lam == 0:
            return
        u = self.vars["obj"] if u is None else u
        self._biharmonic(u, add=grad_obj, scale=2.0 * lam / self._obj_size(), out=grad_obj)

    def hessian_laplacian(self, dobj1, e=None, out=None):
        """2 lam/obj_size Re<dobj1, (lap)^2 e> (rec_mpi.py:445-469) as a
        float64 device scalar (added into out if given)."""
        lam = getattr(self, "lam_laplacian", 0)
        out = self._scalar() if out is None else out
        if lam == 0:
            return out
        e = self.etas["obj"] if e is None else e
        acc = self._redot_dev(dobj1, self._biharmonic(e), self._scalar())
        out += 2.0 * lam / self._obj_size() * acc
        return out

    def _lap_energy_dev(self, u=None):
        """lam/obj_size ||lap(u)||^2 over u's own slices (rec_mpi.py:471-494)."""
        lam = getattr(self, "lam_laplacian", 0)
        if lam == 0:
            return self._scalar()
        u = self.vars["obj"] if u is None else u
        L = self._lap(u, 0, u.shape[0])
        return self._redot_dev(L, L, self._scalar()) * (lam / self._obj_size())

    def _obj_size(self):
        return self.nzobj * self.nobj ** 2

    # ------------------------------------------------------------------ probe fit, device forms
    def _hessian_prbfit_dev(self, prb, dprb1, dprb2, out):
        if self.lam_prbfit == 0:
            return out
        ref = _lib.as_device(self.ref, torch.float32, self.device)
        acc = self._scalar()
        for j in range(self.ndist):
            Dp = self.cl_prop.D(prb[j:j + 1], j)
            D1 = self.cl_prop.D(dprb1[j:j + 1], j)
            D2 = self.cl_prop.D(dprb2[j:j + 1], j)
            acc += self._reduce(2, Dp, ref[j:j + 1], D1, D2)
        out += (2.0 * self.lam_prbfit / self.prb_size) * acc
        return out

    def _prbfit_energy_dev(self, prb):
        out = self._scalar()
        if self.lam_prbfit == 0:
            return out
        ref = _lib.as_device(self.ref, torch.float32, self.device)
        for j in range(self.ndist):
            out += self._reduce(0, self.cl_prop.D(prb[j:j + 1], j), ref[j:j + 1])
        return out * (self.lam_prbfit / self.prb_size)

    # ------------------------------------------------------------------ drivers (rec_mpi.py:235-325)
    @staticmethod
    def _sl(group, st, B):
        return {"prb": group["prb"], "proj": group["proj"][st:st + B], "pos": group["pos"][st:st + B]}

    def gradients(self, vars, grads):
        """Full gradient (rec_mpi.py:305-325): the data term chunk by chunk of
        angles through the fused pipeline, then RT (gF4), the regulariser and
        the probe fit."""
        grads["prb"].zero_()
        grad = self.gradients_fused if self.fused_available() else self.gradients_cascade
        for st, B in self._chunks():
            grad(self._sl(vars, st, B), self._sl(grads, st, B), self.data[st:st + B], st)
        self.gF4(grads["obj"], grads["proj"])
        self.gradient_laplacian(grads["obj"], vars["obj"])
        self.gradient_prbfit(grads["prb"], vars["prb"])

    def hessian_dev(self, vars, grads, etas):
        """hessian (rec_mpi.py:235-248) as a float64 device scalar."""
        out = self._scalar()
        fused = getattr(self, "fused_hessian", True) and self.fused_available()
        for st, B in self._chunks():
            g = self._sl(grads, st, B)
            e = g if grads is etas else self._sl(etas, st, B)
            if fused:
                self.hessian_fused(self._sl(vars, st, B), g, e, self.data[st:st + B], st, out=out)
            else:
                out += self._hessian_cascade_dev(self._sl(vars, st, B), g, e, self.data[st:st + B], st)
        self._hessian_prbfit_dev(vars["prb"], grads["prb"], etas["prb"], out)
        self.hessian_laplacian(grads["obj"], etas["obj"], out=out)
        return out

    def hessian(self, vars, grads, etas):
        """Bilinear form of the full functional (rec_mpi.py:235-248), as a float."""
        return float(self.hessian_dev(vars, grads, etas))

    def hessian_pair_dev(self, vars, grads, etas):
        """(hessian(vars, grads, etas), hessian(vars, etas, etas)) as a float64
        device tensor [2]: the beta step's two forms, whose data terms share
        one fused pass per chunk of angles."""
        if not (getattr(self, "fused_hessian", True) and self.fused_available()):
            return torch.cat([self.hessian_dev(vars, grads, etas), self.hessian_dev(vars, etas, etas)])
        out = torch.zeros(2, dtype=torch.float64, device=self.device)
        for st, B in self._chunks():
            self.hessian_pair_fused(self._sl(vars, st, B), self._sl(grads, st, B), self._sl(etas, st, B),
                                    self.data[st:st + B], st, out=out)
        self._hessian_prbfit_dev(vars["prb"], grads["prb"], etas["prb"], out[0:1])
        self._hessian_prbfit_dev(vars["prb"], etas["prb"], etas["prb"], out[1:2])
        self.hessian_laplacian(grads["obj"], etas["obj"], out=out[0:1])
        self.hessian_laplacian(etas["obj"], etas["obj"], out=out[1:2])
        return out

    def min_dev(self, prb, obj, pos, proj):
        """Objective of Rec.min (rec_mpi.py:815-836) as a float64 device scalar."""
        out = self._scalar()
        for st, B in self._chunks():
            out += self._objective_dev({"prb": prb, "proj": proj[st:st + B], "pos": pos[st:st + B]},
                                       self.data[st:st + B], st)
        out += self._prbfit_energy_dev(prb)
        out += self._lap_energy_dev(obj)
        return out

    def min(self, prb, obj, pos, proj):
        return float(self.min_dev(prb, obj, pos, proj))

    # ------------------------------------------------------------------ logging / checkpoints
    def error_debug(self, vars, i):
        """rec_mpi.py:862-880: objective every err_step iterations, logged and
        appended to self.table (and to path_out/conv.csv if set)."""
        err_step = getattr(self, "err_step", -1)
        if not (err_step != -1 and i % err_step == 0):
            return
        err = self.min(vars["prb"], vars["obj"], vars["pos"], vars["proj"])
        now = time.time()
        ittime = 0.0 if i == -1 else now - self.time_start
        log.warning("iter=%d: %.4fsec err=%1.5e", i, ittime, err)
        self.table.append((i, err, ittime))
        self.time_start = now
        path_out = getattr(self, "path_out", None)
        if path_out:
            os.makedirs(path_out, exist_ok=True)
            with open(os.path.join(path_out, "conv.csv"), "w", newline="") as f:
                w = csv.writer(f)
                w.writerow(["iter", "err", "time"])
                w.writerows(self.table)

    def vis_debug(self, vars, i, writer=None):
        """rec_mpi.py:838-860 without plotting: checkpoint through the writer
        every vis_step iterations."""
        vis_step = getattr(self, "vis_step", -1)
        if not (vis_step != -1 and i % vis_step == 0):
            return
        if writer is not None and i > self.start_iter:
            writer.write_checkpoint(vars, i, self.norm_const)

    # ------------------------------------------------------------------ the loop (rec_mpi.py:138-232)
    def BH(self, writer=None):
        """Bilinear-Hessian conjugate-direction iterations start_iter .. niter-1
        over obj, prb and pos (proj = R obj carried along), as rec_mpi.py:138-232.
        Returns self.vars; obj is normalised by norm_const inside and restored."""
        vars, grads, etas = self.vars, self.grads, self.etas
        self.start_iter = int(getattr(self, "start_iter", 0))
        vars["obj"] /= float(self.norm_const)
        try:
            if self.start_iter == 0:
                mask = self._tomo().mask
                if not np.isscalar(mask):
                    vars["obj"] *= torch.as_tensor(mask, device=self.device)
            self.pos_init = vars["pos"].clone()
            self.fwd_tomo(vars["obj"], out=vars["proj"])
            self.error_debug(vars, -1)
            self.time_start = time.time()
            self.alphas, self.betas = [], []
            timing = bool(getattr(self, "time_phases", False))
            marks = []

            def mark(name):
                # phase boundary: an event on the launch stream (no host wait)
                if timing:
                    e = torch.cuda.Event(enable_timing=True)
                    e.record()
                    marks.append((name, e))

            for i in range(self.start_iter, self.niter):
                mark("start")
                self.gradients(vars, grads)
                mark("gradients")
                self.fwd_tomo(grads["obj"], out=grads["proj"])
                mark("fwd_tomo")
                if i == self.start_iter:
                    beta = self._scalar()
                else:
                    tb = self.hessian_pair_dev(vars, grads, etas)
                    beta = tb[0:1] / tb[1:2]
                mark("hessian_beta")
                ab = torch.cat([beta, self._scalar(-1.0)])
                top = self._scalar()
                for v in ["obj", "pos", "prb"]:
                    self._axpby_redot_dc(etas[v], grads[v], ab, top, -1.0 / self.rho_sq[v])
                self._axpby_dc(etas["proj"], grads["proj"], ab, etas["proj"])
                mark("direction")
                alpha = top / self.hessian_dev(vars, etas, etas)
                mark("hessian_alpha")
                ab = torch.cat([self._scalar(1.0), alpha])
                for v in ["obj", "prb", "pos", "proj"]:
                    self._axpby_dc(vars[v], etas[v], ab, vars[v])
                mark("update")
                self.alphas.append(alpha)
                self.betas.append(beta)
                self.error_debug(vars, i)
                self.vis_debug(vars, i, writer)
            if timing:
                # per iteration, ms spent in each phase (from its previous mark)
                torch.cuda.synchronize()
                self.phase_ms = []
                for name, e in marks:
                    if name == "start":
                        self.phase_ms.append({})
                        prev = e
                        continue
                    self.phase_ms[-1][name] = prev.elapsed_time(e)
                    prev = e
        finally:
            vars["obj"] *= float(self.norm_const)
        return vars

    # ------------------------------------------------------------------ synthetic data (rec_mpi.py:882-902)
    def gen_sqrt_data(self, vars, out=None):
        """out = |F1 F2 F3 (R obj/norm_const, pos, prb)| (rec_mpi.py:882-896)."""
        out = self.data if out is None else out
        obj = vars["obj"] / float(self.norm_const)
        proj = torch.empty((self.ntheta, self.nzobj, self.nobj), dtype=obj.dtype, device=self.device)
        self.fwd_tomo(obj, out=proj)
        for st, B in self._chunks():
            x = self.forward({"prb": vars["prb"], "proj": proj[st:st + B], "pos": vars["pos"][st:st + B]}, st)
            out[st:st + B] = x.abs()
        return out

    def gen_sqrt_ref(self, prb, out=None):
        """out[j] = |D_j prb_j| (rec_mpi.py:898-902)."""
        out = self.ref if out is None else out
        for j in range(self.ndist):
            out[j] = self.cl_prop.D(prb[j:j + 1], j)[0].abs()
        return out
