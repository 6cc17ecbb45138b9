"""Cubic B-spline shift family on gfx950 (drop-in for holotomocupy.shift).

Same constructor and method signatures as the reference class
(src/holotomocupy/shift.py:15-221); device arrays are torch tensors.  Shifts
``r`` are [B, 2] float32 (y, x) and magnifications ``m`` are [B] float32; both
may be strided views (e.g. ``pos[:, j]``) -- they are passed to the kernels
with their stride instead of being copied.
"""
import ctypes

import numpy as np
import torch

from . import _lib

_DTYPES = {"complex64": torch.complex64, "float32": torch.float32}


class Shift:
    """Cubic B-spline shift operator (shift.py:15-48)."""

    def __init__(self, n, npsi, nz, nzpsi, obj_dtype, nchunk=None):
        _lib.require_gpu()
        if str(obj_dtype) not in _DTYPES:
            raise ValueError(f"obj_dtype must be 'complex64' or 'float32', got {obj_dtype!r}")
        self.n, self.npsi, self.nz, self.nzpsi = int(n), int(npsi), int(nz), int(nzpsi)
        self.obj_dtype = str(obj_dtype)
        self.dtype = _DTYPES[self.obj_dtype]
        self.device = torch.device("cuda", torch.cuda.current_device())
        plan = ctypes.c_void_p()
        _lib.call("htamd_shift_plan_create", self.n, self.npsi, self.nz, self.nzpsi,
                  1 if self.dtype == torch.complex64 else 0, ctypes.byref(plan))
        self._plan = plan

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value and _lib._lib is not None:
            _lib._lib.htamd_shift_plan_destroy(plan)
            self._plan = None

    # ------------------------------------------------------------------ helpers
    def _data(self, x, shape_tail):
        x = _lib.as_device(x, self.dtype, self.device)
        if x.ndim == 2:
            x = x[None]
        if x.ndim != 3 or tuple(x.shape[1:]) != shape_tail:
            raise ValueError(f"expected [B, {shape_tail[0]}, {shape_tail[1]}], got {tuple(x.shape)}")
        return x

    def _shifts(self, r, B):
        if not isinstance(r, torch.Tensor):
            r = torch.as_tensor(np.asarray(r, dtype=np.float32))
        r = r.to(device=self.device, dtype=torch.float32)
        r = r.reshape(-1, 2) if r.ndim == 1 else r
        if r.ndim != 2 or r.shape[0] != B or r.shape[1] != 2:
            raise ValueError(f"shifts must be [{B}, 2], got {tuple(r.shape)}")
        if r.stride(1) != 1 or r.stride(0) < 2:
            r = r.contiguous()
        return r, _lib.c_vp(r.data_ptr()), r.stride(0)

    def _mags(self, m, B):
        if not isinstance(m, torch.Tensor):
            m = torch.as_tensor(np.asarray(m, dtype=np.float32))
        m = m.to(device=self.device, dtype=torch.float32)
        if m.ndim == 0:
            m = m.expand(B).contiguous()
        if m.ndim != 1 or m.shape[0] != B:
            raise ValueError(f"magnifications must be [{B}], got {tuple(m.shape)}")
        if m.stride(0) < 1:
            m = m.contiguous()
        return m, _lib.c_vp(m.data_ptr()), m.stride(0)

    # ------------------------------------------------------------------ basis
    def phi(self, t):
        """6 * cubic B-spline (shift.py:54-59), elementwise on a tensor or array."""
        t = torch.as_tensor(t)
        a = t.abs()
        inner = 4 - 6 * t * t + 3 * a * t * t
        outer = (2 - a) ** 3
        return torch.where(a <= 1, inner, torch.where(a < 2, outer, torch.zeros_like(t)))

    # ------------------------------------------------------------------ operators
    def coeff(self, psi):
        """Periodic cubic B-spline prefilter (shift.py:74-82)."""
        x = self._data(psi, (self.nzpsi, self.npsi))
        out = torch.empty_like(x)
        _lib.call("htamd_bspline_prefilter", self._plan, _lib.ptr(x), _lib.ptr(out), x.shape[0],
                  _lib.stream_ptr())
        return out if psi.ndim == 3 else out[0]

    def S(self, c, r, m):
        """out[B, nz, n] = S(c, r, m)  (shift.py:102-111)."""
        c = self._data(c, (self.nzpsi, self.npsi))
        B = c.shape[0]
        r, rp, rs = self._shifts(r, B)
        m, mp, ms = self._mags(m, B)
        out = torch.empty((B, self.nz, self.n), dtype=self.dtype, device=self.device)
        _lib.call("htamd_bspline_interp", self._plan, _lib.ptr(c), rp, rs, mp, ms, _lib.ptr(out), B,
                  _lib.stream_ptr())
        return out

    curlySc = S  # shift.py:159-168 is the same operator on precomputed coefficients

    def Sadj(self, spsi, r, m):
        """out[B, nzpsi, npsi] = S^*(spsi, r, m)  (shift.py:113-122)."""
        g = self._data(spsi, (self.nz, self.n))
        B = g.shape[0]
        r, rp, rs = self._shifts(r, B)
        m, mp, ms = self._mags(m, B)
        out = torch.empty((B, self.nzpsi, self.npsi), dtype=self.dtype, device=self.device)
        _lib.call("htamd_bspline_interp_adj", self._plan, _lib.ptr(g), rp, rs, mp, ms, _lib.ptr(out),
                  B, _lib.stream_ptr())
        return out

    def curlyS(self, psi, r, m):
        """S(coeff(psi))  (shift.py:128-132)."""
        return self.S(self.coeff(psi), r, m)

    def dcurlySc(self, c, r, m, c1, Deltar):
        """First differential of S in (c, r)  (shift.py:170-183)."""
        c = self._data(c, (self.nzpsi, self.npsi))
        c1 = self._data(c1, (self.nzpsi, self.npsi))
        B = c.shape[0]
        r, rp, rs = self._shifts(r, B)
        m, mp, ms = self._mags(m, B)
        dr, dp, ds = self._shifts(Deltar, B)
        out = torch.empty((B, self.nz, self.n), dtype=self.dtype, device=self.device)
        _lib.call("htamd_bspline_d1", self._plan, _lib.ptr(c), _lib.ptr(c1), rp, rs, mp, ms, dp, ds,
                  _lib.ptr(out), B, _lib.stream_ptr())
        return out

    def d2curlySc(self, c, r, m, c1, Deltar1, c2, Deltar2):
        """Second-order term of S  (shift.py:206-221), pairing Deltar1 with c1 as
        the reference kernel does (cuda_kernels.py:437-450)."""
        c = self._data(c, (self.nzpsi, self.npsi))
        c1 = self._data(c1, (self.nzpsi, self.npsi))
        c2 = self._data(c2, (self.nzpsi, self.npsi))
        B = c.shape[0]
        r, rp, rs = self._shifts(r, B)
        m, mp, ms = self._mags(m, B)
        d1, p1, s1 = self._shifts(Deltar1, B)
        d2, p2, s2 = self._shifts(Deltar2, B)
        out = torch.empty((B, self.nz, self.n), dtype=self.dtype, device=self.device)
        _lib.call("htamd_bspline_d2", self._plan, _lib.ptr(c), _lib.ptr(c1), _lib.ptr(c2), rp, rs, mp,
                  ms, p1, s1, p2, s2, _lib.ptr(out), B, _lib.stream_ptr())
        return out

    def dcurlySadjc(self, c, r, m, Deltaphi):
        """Adjoint of the differential (shift.py:185-204): [out1, out2] with
        out1 [B, nzpsi, npsi] and out2 [B, 2] float32."""
        c = self._data(c, (self.nzpsi, self.npsi))
        g = self._data(Deltaphi, (self.nz, self.n))
        B = c.shape[0]
        r, rp, rs = self._shifts(r, B)
        m, mp, ms = self._mags(m, B)
        out1 = torch.empty((B, self.nzpsi, self.npsi), dtype=self.dtype, device=self.device)
        out2 = torch.empty((B, 2), dtype=torch.float32, device=self.device)
        nbytes = ctypes.c_int64()
        _lib.call("htamd_bspline_d1_adj_workspace_size", self._plan, B, ctypes.byref(nbytes))
        ws = torch.empty(max(1, nbytes.value), dtype=torch.uint8, device=self.device)
        _lib.call("htamd_bspline_d1_adj", self._plan, _lib.ptr(c), _lib.ptr(g), rp, rs, mp, ms,
                  _lib.ptr(out1), _lib.ptr(out2), _lib.ptr(ws), nbytes.value, B, _lib.stream_ptr())
        return [out1, out2]

    # ------------------------------------------------------------------ back-shift (initial guess)
    def Sback(self, spsi, r, m):
        """Magnified back-shift of detector images onto the object grid with the
        stretched spline phi(t/m) (shift.py:138-149, sback kernel
        cuda_kernels.py:294-368); complex64 out, as in the reference."""
        g = _lib.as_device(spsi, torch.complex64, self.device)
        g = g[None] if g.ndim == 2 else g
        if g.ndim != 3 or tuple(g.shape[1:]) != (self.nz, self.n):
            raise ValueError(f"expected [B, {self.nz}, {self.n}], got {tuple(spsi.shape)}")
        B = g.shape[0]
        r, rp, rs = self._shifts(r, B)
        m, mp, ms = self._mags(m, B)
        out = torch.empty((B, self.nzpsi, self.npsi), dtype=torch.complex64, device=self.device)
        _lib.call("htamd_bspline_sback", self._plan, _lib.ptr(g), rp, rs, mp, ms, _lib.ptr(out), B,
                  _lib.stream_ptr())
        return out

    def coeffback(self, psi, m):
        """IFFT2(FFT2(psi) / (D_y (x) D_x)) with D(f) = phi(0) + 2 sum_{k=1..4}
        phi(k/m) cos(2 pi k f) (shift.py:84-96); m is one magnification."""
        mv = np.asarray(m.cpu() if isinstance(m, torch.Tensor) else m, dtype=np.float64).reshape(-1)
        if mv.size != 1:
            raise ValueError("coeffback takes one magnification")
        x = _lib.as_device(psi, torch.complex64, self.device)
        x = x[None] if x.ndim == 2 else x
        if x.ndim != 3 or tuple(x.shape[1:]) != (self.nzpsi, self.npsi):
            raise ValueError(f"expected [B, {self.nzpsi}, {self.npsi}], got {tuple(psi.shape)}")
        out = torch.empty_like(x)
        _lib.call("htamd_bspline_coeffback", self._plan, _lib.ptr(x), _lib.ptr(out), x.shape[0], float(mv[0]),
                  _lib.stream_ptr())
        return out

    def curlySback(self, psi, r, m):
        """coeffback(Sback(psi, r, m), m)  (shift.py:151-153)."""
        return self.coeffback(self.Sback(psi, r, m), m)
