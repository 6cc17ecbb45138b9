"""Chunked out-of-core processing on one GPU (drop-in for holotomocupy.chunking).

Same contract as the reference (src/holotomocupy/chunking.py:6-290):
``Chunking(nbytes, chunk).gpu_batch(axis_out, axis_inp, nout, inp_pad)``
decorates ``f(self, *outs, *inps)``; host (NumPy) arguments whose length along
the chunking axis equals the problem size are streamed through device buffers
chunk by chunk, device (torch) arguments are sliced in place, and everything
else is passed whole (NumPy non-chunked inputs are copied to the device once).
If no NumPy array is among the arguments, f runs directly on the device
arrays (the reference's "direct" mode, chunking.py:40-42).

Pipeline: three HIP streams (default priority) carry host-to-device copies of
chunk k, the computation on chunk k-1 and device-to-host copies of chunk k-2
at once, through page-locked staging buffers (double-buffered, reused across
calls) and device buffers carved from one pool of ``nbytes``.  The stages are
ordered by events (a stage waits only for the event that frees the buffer it
overwrites), where the reference synchronises all three streams once per step.

The BLAS-1 helpers (redot_batch, linear_batch, linear_redot_batch,
mulc_batch) run the library's deterministic kernels (htamd_redot,
htamd_axpby, htamd_axpby_redot) on device arrays or on the chunks.
"""
import ctypes

import numpy as np
import torch

from . import _lib

_ALIGN = 256


def _is_host(a):
    return isinstance(a, np.ndarray)


def _is_dev(a):
    return isinstance(a, torch.Tensor)


def _torch_dtype(np_dtype):
    return torch.from_numpy(np.zeros(0, dtype=np_dtype)).dtype


def _slices(axis, sl, ndim):
    res = [slice(None)] * ndim
    res[axis] = sl
    return tuple(res)


def _floats(t):
    """Flat float32 view of a float32 / complex64 device tensor (no copy when
    the tensor is contiguous)."""
    if t.dtype == torch.complex64:
        t = torch.view_as_real(t)
    if t.dtype != torch.float32:
        raise TypeError(f"expected float32 or complex64, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError("BLAS-1 helpers need contiguous device arrays")
    return t


class Chunking:
    """Chunked processing over one axis (reference: chunking.py:6-217)."""

    def __init__(self, nbytes, chunk):
        _lib.require_gpu()
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.gpu_mem = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
        self.stream = [torch.cuda.Stream(device=self.device) for _ in range(3)]
        self.chunk = int(chunk)
        self._pinned = {}
        nb = ctypes.c_int64()
        _lib.call("htamd_blas1_workspace_size", ctypes.byref(nb))
        self._ws = torch.empty(nb.value, dtype=torch.uint8, device=self.device)
        self._res = torch.zeros(1, dtype=torch.float64, device=self.device)

    # ------------------------------------------------------------------ decorator
    def gpu_batch(self, axis_out=0, axis_inp=0, nout=1, inp_pad=0):
        """Decorator with the argument classification of chunking.py:12-88."""

        def decorator(func):
            def inner(*args):
                if not any(_is_host(a) for a in args):
                    func(*args)
                    return
                cl = args[0]
                out = list(args[1:1 + nout])
                inp = list(args[1 + nout:])
                size = inp[0].shape[axis_inp] - inp_pad
                proper_out = nonproper_out = 0
                for o in out:
                    if (_is_host(o) or _is_dev(o)) and o.ndim > axis_out + 1 and o.shape[axis_out] == size:
                        proper_out += 1
                    elif _is_dev(o):
                        nonproper_out += 1
                first = 1 if inp_pad > 0 else 0
                proper_inp, nonproper_inp = first, 0
                for x in inp[first:]:
                    if (_is_host(x) or _is_dev(x)) and x.ndim > axis_inp + 1 and x.shape[axis_inp] == size:
                        proper_inp += 1
                    elif _is_host(x) or _is_dev(x):
                        nonproper_inp += 1
                if int(np.prod(out[0].shape)) == 0:
                    return
                self.run(cl, out, inp, proper_inp, nonproper_inp, proper_out, nonproper_out, axis_out,
                         axis_inp, func, inp_pad)

            return inner

        return decorator

    # ------------------------------------------------------------------ pipeline
    def _pool_views(self, arrs, axis, offset, chunk):
        """Two device buffers per array (chunk rows along axis) from the pool."""
        bufs = [[], []]
        for j in (0, 1):
            for a in arrs:
                shape = list(a.shape)
                shape[axis] = chunk
                dt = a.dtype if _is_dev(a) else _torch_dtype(a.dtype)
                nbytes = int(np.prod(shape)) * torch.empty(0, dtype=dt).element_size()
                if offset + nbytes > self.gpu_mem.numel():
                    raise RuntimeError("Failed to allocate GPU buffers: the Chunking pool is too small")
                bufs[j].append(self.gpu_mem[offset:offset + nbytes].view(dt).view(shape))
                offset = (offset + nbytes + _ALIGN - 1) // _ALIGN * _ALIGN
        return bufs, offset

    def _staging(self, key, shape, dt):
        """Page-locked host buffer reused across calls."""
        k = (key, tuple(shape), dt)
        if k not in self._pinned:
            self._pinned[k] = torch.empty(shape, dtype=dt, pin_memory=True)
        return self._pinned[k]

    def run(self, cl, out, inp, proper_inp, nonproper_inp, proper_out, nonproper_out, axis_out, axis_inp,
            func, inp_pad=0):
        """Chunks through three streams (reference: chunking.py:90-200)."""
        size = inp[0].shape[axis_inp] - inp_pad
        nchunk = int(np.ceil(size / self.chunk))
        out_gpu, off = self._pool_views(out[:proper_out], axis_out, 0, self.chunk)
        if inp_pad > 0 and proper_inp > 0:
            pad_gpu, off = self._pool_views(inp[:1], axis_inp, off, self.chunk + inp_pad)
            rest_gpu, off = self._pool_views(inp[1:proper_inp], axis_inp, off, self.chunk)
            inp_gpu = [[pad_gpu[j][0]] + rest_gpu[j] for j in (0, 1)]
        else:
            inp_gpu, off = self._pool_views(inp[:proper_inp], axis_inp, off, self.chunk)
        for k in range(proper_inp, proper_inp + nonproper_inp):
            if _is_host(inp[k]):
                inp[k] = torch.from_numpy(np.ascontiguousarray(inp[k])).to(self.device)

        pend = {}  # (buffer, out index) -> (host slice, staged pinned view) awaiting the step's sync

        def p2g(buf, k):
            st, end = k * self.chunk, min(size, (k + 1) * self.chunk)
            for j in range(proper_inp):
                extra = inp_pad if (j == 0 and inp_pad > 0) else 0
                nd = inp[j].ndim
                src = _slices(axis_inp, slice(st, end + extra), nd)
                dst = inp_gpu[buf][j][_slices(axis_inp, slice(0, end - st + extra), nd)]
                if _is_dev(inp[j]):
                    dst.copy_(inp[j][src], non_blocking=True)
                else:
                    stage = self._staging(("in", buf, j), inp_gpu[buf][j].shape, inp_gpu[buf][j].dtype)
                    sv = stage[_slices(axis_inp, slice(0, end - st + extra), nd)]
                    np.copyto(sv.numpy(), inp[j][src])
                    dst.copy_(sv, non_blocking=True)

        def compute(buf, k):
            st, end = k * self.chunk, min(size, (k + 1) * self.chunk)
            n = end - st
            ic = []
            for j in range(proper_inp):
                extra = inp_pad if (j == 0 and inp_pad > 0) else 0
                ic.append(inp_gpu[buf][j][_slices(axis_inp, slice(0, n + extra), inp_gpu[buf][j].ndim)])
            oc = [b[_slices(axis_out, slice(0, n), b.ndim)] for b in out_gpu[buf]]
            func(cl, *oc, *out[proper_out:], *ic, *inp[proper_inp:proper_inp + nonproper_inp],
                 *inp[proper_inp + nonproper_inp:])

        def g2p(buf, k):
            st, end = k * self.chunk, min(size, (k + 1) * self.chunk)
            for j in range(proper_out):
                nd = out[j].ndim
                src = out_gpu[buf][j][_slices(axis_out, slice(0, end - st), nd)]
                dst = _slices(axis_out, slice(st, end), nd)
                if _is_dev(out[j]):
                    out[j][dst].copy_(src, non_blocking=True)
                else:
                    stage = self._staging(("out", buf, j), out_gpu[buf][j].shape, out_gpu[buf][j].dtype)
                    sv = stage[_slices(axis_out, slice(0, end - st), nd)]
                    sv.copy_(src, non_blocking=True)
                    pend[(buf, j)] = (dst, sv)

        # Event-ordered pipeline: stream 0 copies chunks in, stream 1 computes,
        # stream 2 copies results out.  Buffer j = k % 2 of chunk k is reused by
        # chunk k + 2, so each stage waits only for the events that free what
        # it overwrites; the host blocks only before refilling a page-locked
        # input buffer (its copy of chunk k - 2 must have left) and before
        # draining an output buffer (its chunk must have arrived).
        h2d, comp, d2h = self.stream
        cur = torch.cuda.current_stream(self.device)
        for s_ in self.stream:
            s_.wait_stream(cur)  # inputs produced on the caller's stream
        ev_in = [torch.cuda.Event() for _ in range(nchunk)]
        ev_comp = [torch.cuda.Event() for _ in range(nchunk)]
        ev_out = [torch.cuda.Event() for _ in range(nchunk)]

        def drain(k):
            ev_out[k].synchronize()
            for (buf, j), (dst, sv) in list(pend.items()):
                if buf == k % 2:
                    out[j][dst] = sv.numpy()
                    del pend[(buf, j)]

        for k in range(nchunk + 3):
            if k < nchunk:
                if k >= 2:
                    ev_in[k - 2].synchronize()  # page-locked input buffer k % 2 is free
                with torch.cuda.stream(h2d):
                    if k >= 2:
                        h2d.wait_event(ev_comp[k - 2])  # device input buffer k % 2 is free
                    p2g(k % 2, k)
                    ev_in[k].record(h2d)
            if 0 <= k - 1 < nchunk:
                with torch.cuda.stream(comp):
                    comp.wait_event(ev_in[k - 1])
                    if k - 1 >= 2:
                        comp.wait_event(ev_out[k - 3])  # device output buffer is free
                    compute((k - 1) % 2, k - 1)
                    ev_comp[k - 1].record(comp)
            if 0 <= k - 2 < nchunk:
                with torch.cuda.stream(d2h):
                    d2h.wait_event(ev_comp[k - 2])
                    g2p((k - 2) % 2, k - 2)
                    ev_out[k - 2].record(d2h)
            if 0 <= k - 3 < nchunk:
                drain(k - 3)  # also frees the page-locked output buffer for chunk k - 1
        for s_ in self.stream:
            cur.wait_stream(s_)

    # ------------------------------------------------------------------ BLAS-1
    def _dev_redot(self, x, y, accumulate=False):
        xf, yf = _floats(x), _floats(y)
        if xf.numel() != yf.numel():
            raise ValueError("redot: size mismatch")
        _lib.call("htamd_redot", _lib.ptr(xf), _lib.ptr(yf), xf.numel(), _lib.ptr(self._res), int(accumulate),
                  _lib.ptr(self._ws), self._ws.numel(), _lib.stream_ptr())

    def _dev_axpby(self, out, x, y, a, b):
        xf, of = _floats(x), _floats(out)
        yf = _floats(y) if y is not None else None
        _lib.call("htamd_axpby", _lib.ptr(xf), _lib.ptr(yf) if yf is not None else None, float(a), float(b),
                  _lib.ptr(of), xf.numel(), _lib.stream_ptr())

    def redot_batch(self, x, y, nout=1):
        """res = Re<x, y>  (chunking.py:234-247)."""
        if _is_dev(x):
            self._dev_redot(x, y)
            return float(self._res.item())
        acc = {"first": True}

        @self.gpu_batch(axis_out=0, axis_inp=0)
        def _redot(self, res, x, y):
            self._dev_redot(x, y, accumulate=not acc["first"])
            acc["first"] = False

        _redot(self, self._res, x, y)
        return float(self._res.item())

    def linear_batch(self, x, y, a, b, out=None):
        """out = a x + b y  (chunking.py:249-262); out defaults to x."""
        if out is None:
            out = x
        if _is_dev(x):
            self._dev_axpby(out, x, y, a, b)
            return

        @self.gpu_batch(axis_out=0, axis_inp=0, nout=1)
        def _linear(self, out, x, y, a, b):
            self._dev_axpby(out, x, y, a, b)

        _linear(self, out, x, y, a, b)

    def linear_redot_batch(self, x, y, a, b):
        """x = a x + b y, returns Re<y, x_new> in one pass  (chunking.py:264-278)."""
        if _is_dev(x):
            xf, yf = _floats(x), _floats(y)
            _lib.call("htamd_axpby_redot", _lib.ptr(xf), _lib.ptr(yf), float(a), float(b), xf.numel(),
                      _lib.ptr(self._res), 0, _lib.ptr(self._ws), self._ws.numel(), _lib.stream_ptr())
            return float(self._res.item())
        acc = {"first": True}

        @self.gpu_batch(axis_out=0, axis_inp=0, nout=2)
        def _linear_redot(self, out, res, x, y, a, b):
            out.copy_(x)
            of, yf = _floats(out), _floats(y)
            _lib.call("htamd_axpby_redot", _lib.ptr(of), _lib.ptr(yf), float(a), float(b), of.numel(),
                      _lib.ptr(self._res), int(not acc["first"]), _lib.ptr(self._ws), self._ws.numel(),
                      _lib.stream_ptr())
            acc["first"] = False

        _linear_redot(self, x, self._res, x, y, a, b)
        return float(self._res.item())

    def mulc_batch(self, out, x, a):
        """out = a x  (chunking.py:280-290)."""
        if _is_dev(x):
            self._dev_axpby(out, x, None, a, 0.0)
            return

        @self.gpu_batch(axis_out=0, axis_inp=0, nout=1)
        def _mulc(self, out, x, a):
            self._dev_axpby(out, x, None, a, 0.0)

        _mulc(self, out, x, a)
