"""Fresnel propagator D / DT on gfx950 (drop-in for holotomocupy.propagation).

Same constructor and method signatures as the reference class
(src/holotomocupy/propagation.py:11-112); device arrays are torch tensors.
The operator runs as two LDS-resident HIP passes (rows, then columns) behind
the C ABI ``htamd_prop_apply``; the padded work buffer of the reference
(``_buf_big``, propagation.py:31) does not exist here.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib


class Propagation:
    """Functionality for Propagation  (reference: propagation.py:11-44)."""

    def __init__(self, n, nz, ntheta, ndist, wavelength, voxelsize, distance):
        _lib.require_gpu()
        self.n = int(n)
        self.nz = int(nz)
        self._ntheta = int(ntheta)
        self.ndist = int(ndist)
        self.wavelength = float(wavelength)
        self.voxelsize = float(voxelsize)
        self.distance = np.ascontiguousarray(np.asarray(distance, dtype=np.float64).reshape(-1)[: self.ndist])
        if len(self.distance) != self.ndist:
            raise ValueError(f"need {self.ndist} distances, got {len(self.distance)}")
        self.device = torch.device("cuda", torch.cuda.current_device())
        plan = ctypes.c_void_p()
        _lib.call("htamd_prop_plan_create", self.n, self.nz, self.ndist, self.wavelength,
                  self.voxelsize, self.distance.ctypes.data_as(_lib.c_dp), ctypes.byref(plan))
        self._plan = plan
        self._fker = None

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value and _lib._lib is not None:
            _lib._lib.htamd_prop_plan_destroy(plan)
            self._plan = None

    @property
    def fker(self):
        """Transfer functions [ndist, 2nz, 2n] (complex64), built on demand in FP64
        exactly as propagation.py:20-28 defines them; the operators do not use it."""
        if self._fker is None:
            fx = torch.fft.fftfreq(2 * self.n, d=self.voxelsize, dtype=torch.float64, device=self.device)
            fy = torch.fft.fftfreq(2 * self.nz, d=self.voxelsize, dtype=torch.float64, device=self.device)
            f2 = fx[None, :] ** 2 + fy[:, None] ** 2
            d = torch.as_tensor(self.distance, device=self.device)[:, None, None]
            ph = -math.pi * self.wavelength * d * f2
            self._fker = (torch.polar(torch.ones_like(ph), ph) / (4 * self.n * self.nz)).to(torch.complex64)
        return self._fker

    def _apply(self, psi, j, adjoint):
        added = psi.ndim == 2
        x = _lib.as_device(psi, torch.complex64, self.device)
        if added:
            x = x[None]
        if x.ndim != 3 or tuple(x.shape[1:]) != (self.nz, self.n):
            raise ValueError(f"expected [B, {self.nz}, {self.n}] or [{self.nz}, {self.n}], got {tuple(psi.shape)}")
        j = int(j)
        if j < 0:
            j += self.ndist
        out = torch.empty_like(x)
        _lib.call("htamd_prop_apply", self._plan, _lib.ptr(x), _lib.ptr(out), x.shape[0], j,
                  1 if adjoint else 0, _lib.stream_ptr())
        return out[0] if added else out

    def D(self, psi, j):
        """Forward propagator  (propagation.py:70-89): crop(IFFT2(FFT2(sympad psi) fker_j))."""
        return self._apply(psi, j, False)

    def DT(self, big_psi, j):
        """Adjoint propagator  (propagation.py:91-112)."""
        return self._apply(big_psi, j, True)

    def apply_into(self, src, dst, j, adjoint=False):
        """dst[b] = D_j(src[b]) (or DT_j) for [B, nz, n] views whose last two
        dimensions are contiguous (e.g. x[:, j] of a [B, ndist, nz, n] array);
        no copies are made.  src may be dst."""
        for t in (src, dst):
            if (t.dtype != torch.complex64 or t.ndim != 3 or tuple(t.shape[1:]) != (self.nz, self.n)
                    or t.stride(2) != 1 or t.stride(1) != self.n or t.device != self.device):
                raise ValueError("apply_into needs complex64 [B, nz, n] views with contiguous rows")
        if src.shape[0] != dst.shape[0]:
            raise ValueError("batch mismatch")
        j = int(j) + (self.ndist if int(j) < 0 else 0)
        _lib.call("htamd_prop_apply_strided", self._plan, _lib.ptr(src), src.stride(0), _lib.ptr(dst),
                  dst.stride(0), src.shape[0], j, 1 if adjoint else 0, _lib.stream_ptr())
        return dst
