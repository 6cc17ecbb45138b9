"""The BH cascade of holotomocupy's Rec on gfx950 (drop-in for the hot part of
holotomocupy.rec_mpi.Rec).

Same method names and signatures as the reference cascade
(src/holotomocupy/rec_mpi.py:517-813): F0..F3, dF0..dF3, d2F_dF0..3, gF0..3,
``gradients_cascade`` (the "BH gradient eval", rec_mpi.py:327-353) and
``hessian_cascade`` (the BH bilinear form, rec_mpi.py:250-303).  Every
operator is a HIP kernel behind the C ABI; device arrays are torch tensors:

  vars / grads / etas = {'prb': [ndist, nz, n] c64, 'proj': [B, nzobj, nobj] obj_dtype,
                         'pos': [B, ndist, 2] f32}
  data = sqrt(intensity) [B, ndist, nz, n] f32

This is the per-operator ("plain") path: it keeps the reference's call
structure (F3 recomputed inside gF0, gF1 and gF2; one D / DT launch per
distance; intermediates [B, ndist, nz, n] materialised).  The fused
multi-pass gradient lives in ``holotomocupy_amd.bh``.
"""
import ctypes
from types import SimpleNamespace

import numpy as np
import torch

from . import _lib
from .bhloop import BHLoop
from .propagation import Propagation
from .shift import Shift


def physics(energy, z1, focustodetectordistance, detector_pixelsize):
    """Wavelength, distances and voxel size of the holotomography geometry
    (rec_mpi.py:55-60)."""
    z1 = np.asarray(z1, dtype=np.float64)
    wavelength = 1.24e-09 / energy
    z2 = focustodetectordistance - z1
    magnifications = focustodetectordistance / z1
    norm_magnifications = magnifications / magnifications[0]
    distance = (z1 * z2) / focustodetectordistance * norm_magnifications ** 2
    voxelsize = detector_pixelsize / magnifications[0]
    return SimpleNamespace(wavelength=wavelength, distance=distance, voxelsize=voxelsize,
                           magnifications=magnifications, norm_magnifications=norm_magnifications)


class Rec(BHLoop):
    """rec_mpi.Rec on one GPU (rec_mpi.py:21-103 construction, 138-325 BH loop,
    250-813 cascade; the loop itself is holotomocupy_amd.bhloop.BHLoop).

    ``args`` carries the reference's fields: energy, detector_pixelsize,
    focustodetectordistance, z1, ndist, ntheta, nz, n, nzobj, nobj, obj_dtype,
    rho, and optionally lam_prbfit (default 0), lam_laplacian (default 0),
    shrink (per-angle, per-distance magnification change, default 0), nchunk
    (angles per launch, default ntheta), and for the BH loop theta, mask_r,
    niter, start_iter, err_step, vis_step, path_out.  vars / grads / etas /
    data / ref are allocated on the device when first used (alloc_arrays).
    """

    _LAZY = ("vars", "grads", "etas", "data", "ref", "table")

    def __getattr__(self, name):
        # only reached when normal lookup fails: allocate the loop's arrays once
        if name in Rec._LAZY and not self.__dict__.get("_allocating", False):
            self.__dict__["_allocating"] = True
            try:
                self.alloc_arrays()
            finally:
                self.__dict__["_allocating"] = False
            return self.__dict__[name]
        raise AttributeError(name)

    def __init__(self, args):
        _lib.require_gpu()
        for key, value in vars(args).items():
            setattr(self, key, value)
        if not hasattr(self, "lam_prbfit"):
            self.lam_prbfit = 0
        self.F = [self.F0, self.F1, self.F2, self.F3]
        self.gF = [self.gF0, self.gF1, self.gF2, self.gF3]
        self.dF = [self.dF0, self.dF1, self.dF2, self.dF3]
        self.d2F_dF = [self.d2F_dF0, self.d2F_dF1, self.d2F_dF2, self.d2F_dF3]
        self.device = torch.device("cuda", torch.cuda.current_device())
        ph = physics(self.energy, self.z1, self.focustodetectordistance, self.detector_pixelsize)
        self.wavelength, self.distance, self.voxelsize = ph.wavelength, ph.distance, ph.voxelsize
        self.norm_magnifications = ph.norm_magnifications
        self.rho_sq = {"obj": args.rho[0] ** 2, "prb": args.rho[1] ** 2, "pos": args.rho[2] ** 2}
        nchunk = getattr(self, "nchunk", self.ntheta)
        self.cl_prop = Propagation(self.n, self.nz, nchunk, self.ndist, self.wavelength, self.voxelsize,
                                   self.distance)
        self.cl_shift = Shift(self.n, self.nobj, self.nz, self.nzobj, self.obj_dtype, nchunk)
        shrink = np.zeros((self.ntheta, self.ndist)) if not hasattr(self, "shrink") else np.asarray(self.shrink)
        # eff_demagnifications (rec_mpi.py:146)
        self.eff_demagnifications = torch.as_tensor(
            ((1 + shrink) / self.norm_magnifications[None, :]).astype(np.float32), device=self.device)
        self.norm_const = np.float32(np.sqrt(self.nobj / self.ntheta))
        self.data_size = self.ntheta * self.ndist * self.nz * self.n
        self.prb_size = self.ndist * self.nz * self.n
        self.obj_real = self.obj_dtype == "float32"
        self._mag = self.eff_demagnifications  # per-chunk view, see _chunk
        self._rws = torch.empty(1024, dtype=torch.float64, device=self.device)

    # ------------------------------------------------------------------ kernels
    def _npix(self):
        return self.nz * self.n

    def _combo(self, terms, B, scale=1.0):
        out = torch.empty((B, self.ndist, self.nz, self.n), dtype=torch.complex64, device=self.device)
        keep = [t.contiguous() for pq in terms for t in pq]  # alive until the launch is enqueued
        ptrs = [_lib.ptr(t) for t in keep]
        while len(ptrs) < 8:
            ptrs.append(None)
        _lib.call("htamd_cas_probe_combo", _lib.ptr(out), *ptrs, B, self.ndist, self._npix(), float(scale),
                  _lib.stream_ptr())
        return out

    def _propagate(self, y0, adjoint=False, out=None):
        out = y0 if out is None else out
        for j in range(self.ndist):
            self.cl_prop.apply_into(y0[:, j], out[:, j], j, adjoint)
        return out

    def _expi(self, mode, x22, y=None, z=None, w=None, want2=False, real_out=False):
        xc = x22.is_complex()
        shape = x22.shape
        out1 = torch.empty(shape, dtype=torch.float32 if real_out else torch.complex64, device=self.device)
        out2 = torch.empty(shape, dtype=torch.complex64, device=self.device) if want2 else None
        args = [x22, y, z, w]
        for t in args:
            if t is not None:
                assert t.is_contiguous()
        _lib.call("htamd_cas_expi", mode, *[_lib.ptr(t) if t is not None else None for t in args],
                  _lib.ptr(out1), _lib.ptr(out2) if out2 is not None else None, x22.numel(),
                  1 if xc else 0, 1 if real_out else 0, _lib.stream_ptr())
        return (out1, out2) if want2 else out1

    def _reduce(self, kind, x, d, y=None, z=None, w=None):
        res = torch.empty(1, dtype=torch.float64, device=self.device)
        _lib.call("htamd_cas_reduce", kind, _lib.ptr(x), _lib.ptr(d), _lib.ptr(y) if y is not None else None,
                  _lib.ptr(z) if z is not None else None, _lib.ptr(w) if w is not None else None,
                  x.numel(), _lib.ptr(res), _lib.ptr(self._rws), self._rws.numel() * 8, _lib.stream_ptr())
        return res[0]

    # ---- F0(x0) = 1/N || |x0| - d ||^2   (rec_mpi.py:517-569)
    def F0(self, x, d):
        return self._reduce(0, x, d) / self.data_size

    def dF0(self, x, y, d, return_x=False):
        return 2 / self.data_size * self._reduce(1, x, d, y)

    def d2F_dF0(self, x, y, z, w, d):
        return 2 / self.data_size * self._reduce(2, x, d, y, z, w)

    def gF0(self, x, y):
        for i in range(1, 4)[::-1]:
            x = self.F[i](x)
        out = torch.empty_like(x)
        _lib.call("htamd_cas_residual", _lib.ptr(x), _lib.ptr(y), _lib.ptr(out), x.numel(),
                  float(2 / self.data_size), _lib.stream_ptr())
        return out

    # ---- x0 = F1(x11, x12) = D(x11 x12)   (rec_mpi.py:572-643)
    def F1(self, x):
        x11, x12 = x
        return self._propagate(self._combo([(x11, x12)], x12.shape[0]))

    def dF1(self, x, y, return_x=True):
        x11, x12 = x
        y11, y12 = y
        y0 = self._propagate(self._combo([(y11, x12), (x11, y12)], x12.shape[0]))
        if return_x:
            return self.F1(x), y0
        return y0

    def d2F_dF1(self, x, y, z, w):
        x11, x12 = x
        y11, y12 = y
        z11, z12 = z
        w11, w12 = w
        terms = [(y11, y12), (y11, y12)] if y12 is z12 else [(y11, z12), (z11, y12)]
        if w11 is not None:
            terms.append((w11, x12))
        if w12 is not None:
            terms.append((x11, w12))
        return self._propagate(self._combo(terms, x12.shape[0]))

    def gF1(self, x, y):
        y0 = y
        for i in range(2, 4)[::-1]:
            x = self.F[i](x)
        x11, x12 = x
        y12 = torch.empty_like(y0)
        self._propagate(y0, adjoint=True, out=y12)
        y11 = torch.empty((self.ndist, self.nz, self.n), dtype=torch.complex64, device=self.device)
        _lib.call("htamd_cas_probe_grad", _lib.ptr(y12), _lib.ptr(x12), _lib.ptr(x11.contiguous()),
                  _lib.ptr(y11), y12.shape[0], self.ndist, self._npix(), 0, _lib.stream_ptr())
        return y11, y12

    # ---- (x11, x12) = F2(x21, x22) = (x21, e^{i x22})   (rec_mpi.py:645-724)
    def F2(self, x):
        x21, x22 = x
        return x21, self._expi(0, x22)

    def dF2(self, x, y, return_x=True):
        x21, x22 = x
        y21, y22 = y
        x12, y12 = self._expi(1, x22, y22, want2=True)
        return ([x21, x12], [y21, y12]) if return_x else [y21, y12]

    def d2F_dF2(self, x, y, z, w):
        x21, x22 = x
        y21, y22 = y
        z21, z22 = z
        w21, w22 = w
        return [w21, self._expi(2, x22, y22, z22, w22)]

    def gF2(self, x, y):
        y11, y12 = y
        x21, x22 = self.F[3](x)
        return [y11, self._expi(3, x22, y12, real_out=self.obj_real)]

    # ---- (x21, x22) = F3(x31, x32, x33) = (x31, S_x33(x32))   (rec_mpi.py:726-813)
    def _m(self, B):
        return self._mag[:B]

    def _field(self, B):
        return torch.empty((B, self.ndist, self.nz, self.n), dtype=self.cl_shift.dtype, device=self.device)

    def F3(self, x):
        x31, x32, x33 = x
        B = x33.shape[0]
        x22 = self._field(B)
        c = self.cl_shift.coeff(x32)
        m = self._m(B)
        for k in range(self.ndist):
            x22[:, k] = self.cl_shift.curlySc(c, x33[:, k], m[:, k])
        return [x31, x22]

    def dF3(self, x, y, return_x=True):
        x31, x32, x33 = x
        y31, y32, y33 = y
        B = x33.shape[0]
        m = self._m(B)
        c = self.cl_shift.coeff(x32)
        c1 = self.cl_shift.coeff(y32)
        y22 = self._field(B)
        for k in range(self.ndist):
            y22[:, k] = self.cl_shift.dcurlySc(c, x33[:, k], m[:, k], c1, y33[:, k])
        if return_x:
            x22 = self._field(B)
            for k in range(self.ndist):
                x22[:, k] = self.cl_shift.curlySc(c, x33[:, k], m[:, k])
            return [x31, x22], [y31, y22]
        return [y31, y22]

    def d2F_dF3(self, x, y, z, w):
        x31, x32, x33 = x
        y31, y32, y33 = y
        z31, z32, z33 = z
        w31, w32, w33 = w
        B = x33.shape[0]
        m = self._m(B)
        c = self.cl_shift.coeff(x32)
        cy = self.cl_shift.coeff(y32)
        cz = self.cl_shift.coeff(z32)
        y22 = self._field(B)
        for k in range(self.ndist):
            y22[:, k] = self.cl_shift.d2curlySc(c, x33[:, k], m[:, k], cy, y33[:, k], cz, z33[:, k])
        if w32 is not None:
            cw = self.cl_shift.coeff(w32)
            for k in range(self.ndist):
                y22[:, k] += self.cl_shift.dcurlySc(c, x33[:, k], m[:, k], cw, w33[:, k])
        return [w31, y22]

    def gF3(self, x, y):
        y21, y22 = y
        x31, x32, x33 = x
        B = x33.shape[0]
        m = self._m(B)
        y32 = torch.zeros((B, self.nzobj, self.nobj), dtype=self.cl_shift.dtype, device=self.device)
        y33 = torch.empty((B, self.ndist, 2), dtype=torch.float32, device=self.device)
        c = self.cl_shift.coeff(x32)
        for k in range(self.ndist):
            dpsi, dr = self.cl_shift.dcurlySadjc(c, x33[:, k], m[:, k], y22[:, k].contiguous())
            y32 += dpsi
            y33[:, k] = dr
        return [y21, self.cl_shift.coeff(y32), y33]

    # ------------------------------------------------------------------ drivers
    def _chunk(self, st, B):
        self._mag = self.eff_demagnifications[st:st + B]

    def gradients_cascade(self, vars, grads, data=None, st=0):
        """Cascade gradient of the data term for the angles [st, st+B)
        (rec_mpi.py:327-353): writes grads['proj'], grads['pos'] and adds to
        grads['prb'] (which the caller zeroes), each scaled by rho^2."""
        data = self.data if data is None else data
        self._chunk(st, vars["proj"].shape[0])
        x = [vars["prb"], vars["proj"], vars["pos"]]
        y = data
        for i in range(len(self.gF)):
            y = self.gF[i](x, y)
        grads["prb"] += y[0] * self.rho_sq["prb"]
        grads["proj"][:] = y[1] * self.rho_sq["obj"]
        grads["pos"][:] = y[2] * self.rho_sq["pos"]

    def fused_available(self):
        """True when the fused passes take this geometry: power-of-two detector
        sizes up to 2048.  Other sizes (any even size up to 8192) run on the
        per-operator kernels, which the BH loop then uses instead."""
        if getattr(self, "_bh_ok", None) is None:
            from .bh import BHGradient
            try:
                self._bh = BHGradient(self)
                self._bh_ok = True
            except _lib.HtamdError as e:
                if "(-2)" not in str(e):
                    raise
                self._bh_ok = False
        return self._bh_ok

    def gradients_fused(self, vars, grads, data=None, st=0, objective=None):
        """Same result as gradients_cascade through the fused three-pass HIP
        pipeline (holotomocupy_amd.bh); grads['prb'] is accumulated into.
        objective: optional float64 [1] device tensor receiving
        sum (|x0| - d)^2 over these angles."""
        from .bh import BHGradient
        data = self.data if data is None else data
        B = vars["proj"].shape[0]
        if getattr(self, "_bh", None) is None:
            self._bh = BHGradient(self)
        mag = self.eff_demagnifications[st:st + B].contiguous()
        self._bh(vars["prb"], vars["proj"], vars["pos"], data, mag, grads["proj"], grads["pos"],
                 grads["prb"], accumulate_prb=True, objective=objective)

    def hessian_fused(self, vars, grads, etas, data=None, st=0, out=None):
        """hessian_cascade through the fused two-pass HIP pipeline
        (htamd_bh_hessian), as a float64 device tensor [1] (added into out if
        given).  grads is etas selects the equal-direction form."""
        from .bh import BHGradient
        data = self.data if data is None else data
        B = vars["proj"].shape[0]
        if getattr(self, "_bh", None) is None:
            self._bh = BHGradient(self)
        mag = self.eff_demagnifications[st:st + B].contiguous()
        res = torch.zeros(1, dtype=torch.float64, device=self.device) if out is None else out
        z = grads if grads is etas else etas
        self._bh.hessian(vars, grads, z, data, mag, res, accumulate=True)
        return res

    def hessian_pair_fused(self, vars, grads, etas, data=None, st=0, out=None):
        """(H(grads, etas), H(etas, etas)) of the data term in one fused pass
        (htamd_bh_hessian_pair), added into out (float64 device tensor [2],
        zeros if not given) and returned."""
        from .bh import BHGradient
        data = self.data if data is None else data
        B = vars["proj"].shape[0]
        if getattr(self, "_bh", None) is None:
            self._bh = BHGradient(self)
        mag = self.eff_demagnifications[st:st + B].contiguous()
        res = torch.zeros(2, dtype=torch.float64, device=self.device) if out is None else out
        self._bh.hessian_pair(vars, grads, etas, data, mag, res, accumulate=True)
        return res

    def hessian_cascade(self, vars, grads, etas, data=None, st=0):
        """BH bilinear form of the data term (rec_mpi.py:250-303), as a float."""
        return float(self._hessian_cascade_dev(vars, grads, etas, data, st))

    def _hessian_cascade_dev(self, vars, grads, etas, data=None, st=0):
        """hessian_cascade as a float64 device tensor [1] (no host sync)."""
        data = self.data if data is None else data
        self._chunk(st, vars["proj"].shape[0])
        x = [vars["prb"], vars["proj"], vars["pos"]]
        y = [grads["prb"], grads["proj"], grads["pos"]]
        z = y if grads is etas else [etas["prb"], etas["proj"], etas["pos"]]
        y_is_z = y is z
        w = [None, None, None]
        for i in range(1, len(self.F))[::-1]:
            w = self.d2F_dF[i](x, y, z, w)
            fx, y = self.dF[i](x, y)
            z = y if y_is_z else self.dF[i](x, z, return_x=False)
            x = fx
        return self.d2F_dF0(x, y, z, w, data).reshape(1)

    def objective(self, vars, data=None, st=0):
        """Data term of Rec.min (rec_mpi.py:815-835) for the given angles."""
        return float(self._objective_dev(vars, data, st))

    def _objective_dev(self, vars, data=None, st=0):
        data = self.data if data is None else data
        self._chunk(st, vars["proj"].shape[0])
        y = [vars["prb"], vars["proj"], vars["pos"]]
        for i in range(1, len(self.F))[::-1]:
            y = self.F[i](y)
        return self.F0(y, data).reshape(1)

    # ------------------------------------------------------------------ probe fit
    def gradient_prbfit(self, grad_prb, prb):
        """Adds the gradient of lam_prbfit/prb_size || |D prb| - ref ||^2 to
        grad_prb, scaled by rho_prb^2 (rec_mpi.py:365-378); self.ref is the
        [ndist, nz, n] float32 reference amplitude."""
        if self.lam_prbfit == 0:
            return
        ref = _lib.as_device(self.ref, torch.float32, self.device)
        for j in range(self.ndist):
            tmp = self.cl_prop.D(prb[j:j + 1], j)
            res = torch.empty_like(tmp)
            _lib.call("htamd_cas_residual", _lib.ptr(tmp), _lib.ptr(ref[j:j + 1]), _lib.ptr(res), tmp.numel(),
                      float(2 * self.lam_prbfit / self.prb_size), _lib.stream_ptr())
            grad_prb[j:j + 1] += self.cl_prop.DT(res, j) * self.rho_sq["prb"]

    def hessian_prbfit(self, prb, dprb1, dprb2):
        """Bilinear form of the probe-fit term (rec_mpi.py:380-398), as a float."""
        if self.lam_prbfit == 0:
            return 0
        ref = _lib.as_device(self.ref, torch.float32, self.device)
        out = 0.0
        for j in range(self.ndist):
            Dp = self.cl_prop.D(prb[j:j + 1], j)
            D1 = self.cl_prop.D(dprb1[j:j + 1], j)
            D2 = self.cl_prop.D(dprb2[j:j + 1], j)
            out += 2 * float(self._reduce(2, Dp, ref[j:j + 1], D1, D2))
        return self.lam_prbfit * out / self.prb_size

    # ------------------------------------------------------------------ tomography
    def _tomo(self):
        """Tomo(nobj, nchunk, theta, mask_r) of the reconstruction (rec_mpi.py:71-74);
        needs args.theta (and args.mask_r, default 0)."""
        if getattr(self, "cl_tomo", None) is None:
            from .tomo import Tomo
            # the plan does not depend on the slice count: accept whole device slabs
            self.cl_tomo = Tomo(self.nobj, max(getattr(self, "nchunk", self.nzobj), self.nzobj), self.theta,
                                getattr(self, "mask_r", 0))
        return self.cl_tomo

    def _tomo_slab(self):
        """Slices per R / RT call on device-resident arrays: as many as keep the
        Radon workspace (htamd_tomo_workspace_size: 6 nobj^2 + ntheta nobj
        complex64 per slice) within 4 GiB, a multiple of 32 (the RT gather's
        slice block) when that many fit, and at least nchunk."""
        per_slice = 8 * (6 * self.nobj * self.nobj + self.ntheta * self.nobj)
        slab = max(1, (4 << 30) // per_slice)
        if slab >= 32:
            slab -= slab % 32
        return int(min(self.nzobj, max(slab, getattr(self, "nchunk", 1))))

    def _chunking(self):
        """Chunking pool of rec_mpi.py:39-42 (16 nchunk nobj max(nobj, nzobj, ntheta) x 8 bytes)."""
        if getattr(self, "cl_chunk", None) is None:
            from .chunking import Chunking
            nchunk = getattr(self, "nchunk", self.nzobj)
            nbytes = 16 * nchunk * self.nobj * max(self.nobj, self.nzobj, self.ntheta) * 8
            self.cl_chunk = Chunking(nbytes, nchunk)
        return self.cl_chunk

    def fwd_tomo(self, obj, out):
        """out [ntheta, nzobj, nobj] = R(obj [nzobj, nobj, nobj]) chunk by chunk
        over the slices (rec_mpi.py:498-507); host or device arrays."""
        cl = self._chunking()
        tomo = self._tomo()
        if isinstance(obj, torch.Tensor) and obj.is_cuda and isinstance(out, torch.Tensor) and out.is_cuda:
            slab = self._tomo_slab()
            if (slab >= obj.shape[0] and out.is_contiguous() and out.dtype == obj.dtype
                    and obj.dtype in (torch.float32, torch.complex64)):
                return tomo.R(obj, out=out)  # one slab: straight into out
            for z0 in range(0, obj.shape[0], slab):
                z1 = min(obj.shape[0], z0 + slab)
                out[:, z0:z1] = tomo.R(obj[z0:z1])
            return out

        @cl.gpu_batch(axis_out=1, axis_inp=0, nout=1)
        def _fwd_tomo(self, out, obj):
            out[:] = tomo.R(obj)

        _fwd_tomo(self, out, obj)
        return out

    def gF4(self, gradu, gradproj):
        """gradu [nzobj, nobj, nobj] = RT(gradproj [ntheta, nzobj, nobj]) chunk by
        chunk (rec_mpi.py:355-362)."""
        cl = self._chunking()
        tomo = self._tomo()
        if isinstance(gradu, torch.Tensor) and gradu.is_cuda and isinstance(gradproj, torch.Tensor) \
                and gradproj.is_cuda:
            slab = self._tomo_slab()
            for z0 in range(0, gradu.shape[0], slab):
                z1 = min(gradu.shape[0], z0 + slab)
                src = gradproj if (z0, z1) == (0, gradu.shape[0]) else gradproj[:, z0:z1].contiguous()
                dst = gradu[z0:z1]  # whole slices: contiguous when gradu is
                if dst.is_contiguous() and dst.dtype == src.dtype and src.dtype in (torch.float32,
                                                                                     torch.complex64):
                    tomo.RT(src, out=dst)
                else:
                    dst[:] = tomo.RT(src)
            return

        @cl.gpu_batch(axis_out=0, axis_inp=1, nout=1)
        def _gF4(self, gradu, gradproj):
            gradu[:] = tomo.RT(gradproj)

        _gF4(self, gradu, gradproj)

    def forward(self, vars, st=0):
        """F1(F2(F3(x))) [B, ndist, nz, n]; its modulus is the synthetic data
        (rec_mpi.py:882-896)."""
        self._chunk(st, vars["proj"].shape[0])
        y = [vars["prb"], vars["proj"], vars["pos"]]
        for i in range(1, len(self.F))[::-1]:
            y = self.F[i](y)
        return y


_ = ctypes  # ctypes is used through _lib
