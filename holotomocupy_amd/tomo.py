"""USFFT Radon transform R / RT on gfx950 (drop-in for holotomocupy.tomo).

Same constructor and method signatures as the reference class
(src/holotomocupy/tomo.py:9-210); device arrays are torch tensors.  R and RT
run as three hand-written HIP passes each behind the C ABI ``htamd_tomo_R`` /
``htamd_tomo_RT``: the oversampled 2-D FFT as two LDS passes with the sign
modulation, zero-embedding and deapodisation folded in, and the polar gather
(scatter) with the fused 1-D line transform.
"""
import ctypes
import logging

import numpy as np
import torch

from . import _lib

log = logging.getLogger(__name__)


def support_mask(n, mask_r):
    """Smoothed circular support of tomo.py:34-44 (host FP64, rounded to
    float32 as there); 1.0 when mask_r <= 0."""
    if not mask_r > 0:
        return 1.0
    t1d = np.linspace(-1, 1, n)
    x, y = np.meshgrid(t1d, t1d)
    circ = (x ** 2 + y ** 2 < mask_r).astype("float32")
    g = np.exp(-(20 ** 2) * (x ** 2 + y ** 2))
    fcirc = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(circ)))
    fg = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(g)))
    mask = np.fft.fftshift(np.fft.ifft2(np.fft.fftshift(fcirc * fg))).real.astype("float32")
    mask /= np.amax(mask)
    return mask


def redot(a, b):
    """Re<a, b> in FP64 (utils.py redot)."""
    if a.is_complex():
        return float(torch.sum(a.real.double() * b.real.double() + a.imag.double() * b.imag.double()))
    return float(torch.sum(a.double() * b.double()))


class Tomo:
    """Functionality for Radon transforms  (reference: tomo.py:9-54)."""

    def __init__(self, n, nz, theta, mask_r):
        _lib.require_gpu()
        self.n = int(n)
        self._nz = int(nz)
        theta = np.asarray(theta.cpu() if isinstance(theta, torch.Tensor) else theta, dtype=np.float64).reshape(-1)
        self.ntheta = len(theta)
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.theta = torch.as_tensor(theta.astype("float32"), device=self.device)
        self.mask = support_mask(self.n, mask_r)
        mask64 = None
        if not np.isscalar(self.mask):
            mask64 = np.ascontiguousarray(self.mask, dtype=np.float64)
        th = np.ascontiguousarray(theta)
        plan = ctypes.c_void_p()
        _lib.call("htamd_tomo_plan_create", self.n, self.ntheta, th.ctypes.data_as(_lib.c_dp),
                  mask64.ctypes.data_as(_lib.c_dp) if mask64 is not None else None, ctypes.byref(plan))
        self._plan = plan

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value and _lib._lib is not None:
            _lib._lib.htamd_tomo_plan_destroy(plan)
            self._plan = None

    def _workspace(self, nz):
        nbytes = ctypes.c_int64()
        _lib.call("htamd_tomo_workspace_size", self._plan, nz, ctypes.byref(nbytes))
        return torch.empty(nbytes.value, dtype=torch.uint8, device=self.device)

    def _run(self, name, x, out_shape, out=None):
        is_complex = x.is_complex()
        x = _lib.as_device(x, torch.complex64 if is_complex else torch.float32, self.device)
        nz = out_shape[0] if name == "htamd_tomo_RT" else out_shape[1]
        if nz > self._nz:
            raise ValueError(f"{nz} slices exceed the {self._nz} this Tomo was built for")
        if out is None:
            out = torch.empty(out_shape, dtype=x.dtype, device=self.device)
        elif (tuple(out.shape) != tuple(out_shape) or out.dtype != x.dtype or not out.is_contiguous()
              or out.device != self.device):
            raise ValueError(f"out must be a contiguous {x.dtype} tensor of shape {tuple(out_shape)} on {self.device}")
        ws = self._workspace(nz)
        _lib.call(name, self._plan, _lib.ptr(x), 1 if is_complex else 0, _lib.ptr(out), nz, _lib.ptr(ws),
                  ws.numel(), _lib.stream_ptr())
        return out

    def R(self, obj, out=None):
        """Radon transform (tomo.py:56-86): obj [nz', n, n] -> [ntheta, nz', n]
        (float32 in -> float32 out, complex64 -> complex64); written into `out`
        (contiguous, that shape) when given."""
        if obj.ndim != 3 or tuple(obj.shape[1:]) != (self.n, self.n):
            raise ValueError(f"expected [nz, {self.n}, {self.n}], got {tuple(obj.shape)}")
        return self._run("htamd_tomo_R", obj, (self.ntheta, obj.shape[0], self.n), out)

    def RT(self, data, out=None):
        """Adjoint Radon transform (tomo.py:88-116): [ntheta, nz', n] -> [nz', n, n];
        written into `out` (contiguous, that shape) when given."""
        if data.ndim != 3 or data.shape[0] != self.ntheta or data.shape[2] != self.n:
            raise ValueError(f"expected [{self.ntheta}, nz, {self.n}], got {tuple(data.shape)}")
        return self._run("htamd_tomo_RT", data, (data.shape[1], self.n, self.n), out)

    _FILTERS = {"ramp": 0, "shepp": 1, "parzen": 2}

    def _filter_kind(self, filter_name):
        if filter_name not in self._FILTERS:
            raise ValueError(f"Unknown filter '{filter_name}'. Choose: ramp, shepp, parzen.")
        return self._FILTERS[filter_name]

    @staticmethod
    def _sino_dtype(data):
        """complex64 for a complex sinogram, float32 otherwise (host or device)."""
        if isinstance(data, torch.Tensor):
            return torch.complex64 if data.is_complex() else torch.float32
        return torch.complex64 if np.iscomplexobj(data) else torch.float32

    def _filter_sino(self, data, filter_name):
        """Ramp / Shepp-Logan / Parzen filter along the detector axis
        (tomo.py:118-162): IFFT(h FFT(data)) in one LDS pass per line
        (htamd_tomo_filter); same dtype as data."""
        kind = self._filter_kind(filter_name)
        x = _lib.as_device(data, self._sino_dtype(data), self.device)
        if x.ndim != 3 or x.shape[0] != self.ntheta or x.shape[2] != self.n:
            raise ValueError(f"expected [{self.ntheta}, nz, {self.n}], got {tuple(x.shape)}")
        out = torch.empty_like(x)
        _lib.call("htamd_tomo_filter", self._plan, _lib.ptr(x), 1 if x.is_complex() else 0, kind, _lib.ptr(out),
                  x.shape[1], _lib.stream_ptr())
        return out

    def fbp(self, data, filter_name="ramp"):
        """Filtered back-projection (tomo.py:164-186): RT(filter(data)) sqrt(n/ntheta),
        with the filter applied inside RT's own line transform (htamd_tomo_fbp)."""
        kind = self._filter_kind(filter_name)
        x = _lib.as_device(data, self._sino_dtype(data), self.device)
        if x.ndim != 3 or x.shape[0] != self.ntheta or x.shape[2] != self.n:
            raise ValueError(f"expected [{self.ntheta}, nz, {self.n}], got {tuple(x.shape)}")
        nz = x.shape[1]
        if nz > self._nz:
            raise ValueError(f"{nz} slices exceed the {self._nz} this Tomo was built for")
        out = torch.empty((nz, self.n, self.n), dtype=x.dtype, device=self.device)
        ws = self._workspace(nz)
        _lib.call("htamd_tomo_fbp", self._plan, _lib.ptr(x), 1 if x.is_complex() else 0, kind, _lib.ptr(out), nz,
                  _lib.ptr(ws), ws.numel(), _lib.stream_ptr())
        return out

    def rec_tomo(self, d, niter=1):
        """Conjugate-gradient tomography for an initial guess (tomo.py:188-210)."""
        u = torch.zeros((d.shape[1], self.n, self.n), dtype=d.dtype, device=self.device)
        Ru = self.R(u)
        eta = Reta = None
        for k in range(niter):
            if k % 32 == 0:
                log.info("rec_tomo iter %d: err=%.6e", k, float(torch.linalg.norm(Ru - d)) ** 2)
            grad = self.RT(2 * (Ru - d))
            Rgrad = self.R(grad)
            if k == 0:
                eta, Reta = -grad, -Rgrad
            else:
                beta = redot(Rgrad, Reta) / redot(Reta, Reta)
                eta = beta * eta - grad
                Reta = beta * Reta - Rgrad
            alpha = -redot(grad, eta) / (2 * redot(Reta, Reta))
            u += alpha * eta
            Ru += alpha * Reta
        return u
