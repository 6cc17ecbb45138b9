"""Continuous-batching request loop for BH operator solves on one GPU.

Four kinds of requests: BH gradients (the main kind, below), propagations
D / DT of batches of fields (packed per geometry, operator and distance into
one launch), Radon transforms R / RT of slabs (stacked along z per grid, angle
set, mask and operator into one launch) and BH bilinear forms (run by the
fused bilinear-form kernel, one launch sequence per request).  Gradient requests (a chunk of angles of some
reconstruction: probe, projected object, positions, measured amplitudes,
geometry) are
  * bucketed by geometry (detector and object grid, distances, physics), since
    one fused launch needs one propagator and one shift plan;
  * admitted while their in-flight device footprint (result copies, packed
    batches) fits the HBM budget, in arrival order within a bucket;
  * packed into device-resident staging arrays (angles of every request
    contiguous, each request with its own probe slot, its own 2/data_size and
    rho scales), so one fused gradient launch (holotomocupy_amd.bh,
    htamd_bh_gradient) serves the whole batch; staging and workspaces are views
    of one byte arena per stream, shared by all geometries;
  * launched on one of two streams, so packing and unpacking of one batch
    overlap the kernels of the previous one; the fused launch sequence of a
    batch shape (geometry, angles, probes) seen before is captured as a HIP
    graph and replayed.
A request's result does not depend on the batch it shares: every per-angle
quantity is computed by the same arithmetic, and probe gradients are summed
per probe slot in angle order (deterministic kernels, no atomics).

The scheduling core (`Scheduler`) is plain Python and is exercised on the CPU
with a fake executor; `GpuExecutor` runs the HIP kernels.
"""
import collections
import gc
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np


@dataclass(frozen=True)
class Geometry:
    """What one fused launch shares: grids, distances and physics (the fields of
    rec_mpi.Rec's args that fix the operators, rec_mpi.py:55-74)."""
    n: int
    nz: int
    nobj: int
    nzobj: int
    ndist: int
    energy: float = 17.1
    z1: tuple = (5.110e-3, 5.464e-3, 6.879e-3, 9.817e-3)
    focustodetectordistance: float = 1.217
    detector_pixelsize: float = 1.4760147601476e-6 * 2
    obj_dtype: str = "complex64"


@dataclass
class GradientRequest:
    """One BH gradient evaluation (Rec.gradients_cascade over its angles).
    Device tensors: prb [ndist, nz, n] c64, proj [A, nzobj, nobj], pos
    [A, ndist, 2] f32, data [A, ndist, nz, n] f32; data_size is the
    reconstruction's ntheta*ndist*nz*n (rec_mpi.py:90); rho = (obj, prb, pos)."""
    geometry: Geometry
    prb: object
    proj: object
    pos: object
    data: object
    data_size: float
    rho: tuple = (1.0, 1.0, 1.0)
    nangles: int = 0
    future: Future = field(default_factory=Future)
    t_submit: float = 0.0
    t_done: float = 0.0
    nbytes: int = 0

    def __post_init__(self):
        self.nangles = int(self.proj.shape[0])

    @property
    def key(self):
        return self.geometry


@dataclass
class OperatorRequest:
    """D_j or DT_j of a batch of fields (Propagation.D / DT,
    propagation.py:70-112): x [B, nz, n] complex64 device tensor.  Requests of
    one geometry, operator and distance share one launch; the result is the
    request's [B, nz, n] output."""
    geometry: Geometry
    op: str
    j: int
    x: object
    nangles: int = 0
    future: Future = field(default_factory=Future)
    t_submit: float = 0.0
    t_done: float = 0.0
    nbytes: int = 0

    def __post_init__(self):
        if self.op not in ("D", "DT"):
            raise ValueError("op must be 'D' or 'DT'")
        self.nangles = int(self.x.shape[0])

    @property
    def key(self):
        return (self.geometry, "prop", self.op, int(self.j))


@dataclass
class TomoRequest:
    """R or RT of a slab (Tomo.R / RT, tomo.py:56-116) at angles theta (a
    tuple of floats) with support mask radius mask_r: x is obj [nz, n, n] for R,
    data [ntheta, nz, n] for RT (complex64 or float32 device tensors).
    Requests of one grid, angle set, mask and operator share one launch, their
    slabs stacked along z; the result is the request's own output."""
    n: int
    theta: tuple
    mask_r: float
    op: str
    x: object
    nangles: int = 0
    future: Future = field(default_factory=Future)
    t_submit: float = 0.0
    t_done: float = 0.0
    nbytes: int = 0

    def __post_init__(self):
        if self.op not in ("R", "RT"):
            raise ValueError("op must be 'R' or 'RT'")
        self.nangles = int(self.x.shape[0] if self.op == "R" else self.x.shape[1])  # slices

    @property
    def key(self):
        return ("tomo", int(self.n), tuple(self.theta), float(self.mask_r), self.op, self.x.dtype)


@dataclass
class BilinearRequest:
    """The BH bilinear form d^2F(x)(y, z) of the data term over the request's
    angles (Rec.hessian_cascade, rec_mpi.py:250-303): x, y, z are
    {'prb', 'proj', 'pos'} dicts of device tensors (z None: z = y), data
    [A, ndist, nz, n].  The result is a float64 device tensor [1]."""
    geometry: Geometry
    x: dict
    y: dict
    z: object
    data: object
    data_size: float
    nangles: int = 0
    future: Future = field(default_factory=Future)
    t_submit: float = 0.0
    t_done: float = 0.0
    nbytes: int = 0

    def __post_init__(self):
        self.nangles = int(self.x["proj"].shape[0])

    @property
    def key(self):
        return (self.geometry, "hess")


class Scheduler:
    """Admission by HBM budget and FIFO batching per bucket (a geometry for
    gradients; geometry, operator and distance for propagations; geometry for
    bilinear forms).

    executor.footprint(req) -> bytes the request holds while in flight;
    executor.run(geometry, batch) -> handle (enqueues the batch);
    executor.done(handle) -> bool; executor.finish(handle, batch) sets results.
    """

    def __init__(self, executor, hbm_budget, max_batch_angles=128, max_inflight=2):
        self.ex = executor
        self.budget = int(hbm_budget)
        self.max_batch = int(max_batch_angles)
        self.max_inflight = max_inflight
        self.pending = collections.deque()      # submitted, not admitted
        self.admitted = collections.OrderedDict()  # geometry -> deque of admitted requests
        self.inflight = []                      # (handle, geometry, batch)
        self.in_use = 0
        self.rejected = 0

    def submit(self, req):
        prepare = getattr(self.ex, "prepare", None)
        if prepare is not None:
            prepare(req)
        req.nbytes = int(self.ex.footprint(req))
        if req.nbytes > self.budget:
            self.rejected += 1
            req.future.set_exception(MemoryError(
                f"request needs {req.nbytes} bytes of device memory, above the budget of {self.budget}"))
            return req.future
        if req.nangles > self.max_batch:
            self.rejected += 1
            req.future.set_exception(ValueError(f"request of {req.nangles} angles exceeds the batch limit"))
            return req.future
        req.t_submit = time.perf_counter()
        self.pending.append(req)
        return req.future

    def _admit(self):
        while self.pending and self.in_use + self.pending[0].nbytes <= self.budget:
            r = self.pending.popleft()
            self.in_use += r.nbytes
            self.admitted.setdefault(r.key, collections.deque()).append(r)

    def _next_batch(self):
        for geo, q in list(self.admitted.items()):
            if not q:
                del self.admitted[geo]
                continue
            batch, angles = [], 0
            while q and angles + q[0].nangles <= self.max_batch:
                r = q.popleft()
                batch.append(r)
                angles += r.nangles
            # rotate buckets so one geometry cannot starve the others
            self.admitted.move_to_end(geo)
            return geo, batch
        return None, None

    def _release(self, batch, error=None):
        now = time.perf_counter()
        for r in batch:
            r.t_done = now
            self.in_use -= r.nbytes
            f = r.future  # a client callback may have dropped it already
            if error is not None and f is not None and not f.done():
                f.set_exception(error)

    def _retire(self):
        keep = []
        for h, geo, batch in self.inflight:
            try:
                if not self.ex.done(h):
                    keep.append((h, geo, batch))
                    continue
                self.ex.finish(h, batch)
                self._release(batch)
            except Exception as e:  # a failed batch fails its requests, not the loop
                self._release(batch, e)
        self.inflight = keep

    def step(self):
        """One scheduling round; returns True if anything moved."""
        before = (len(self.pending), len(self.inflight), self.in_use)
        self._retire()
        self._admit()
        while len(self.inflight) < self.max_inflight:
            geo, batch = self._next_batch()
            if not batch:
                break
            try:
                self.inflight.append((self.ex.run(geo, batch), geo, batch))
            except Exception as e:
                self._release(batch, e)
        return (len(self.pending), len(self.inflight), self.in_use) != before

    def idle(self):
        return not self.pending and not self.inflight and not any(self.admitted.values())


class GpuExecutor:
    """Runs batches on the GPU: gradients through the fused BH gradient,
    propagations through one packed D / DT launch, bilinear forms through the
    fused bilinear-form kernels.

    Each stream owns one byte arena, shared by every geometry: a stream runs
    its batches in order, so a batch's staging arrays and workspace are views
    at fixed offsets of its stream's arena, and the arena only grows (to the
    largest batch seen).  A gradient batch shape (geometry, angles, probes)
    runs eagerly the first time and is captured as a HIP graph the second
    time, then replayed; growing an arena drops the graphs of its stream."""

    def __init__(self, use_graphs=True, nstreams=2):
        import torch
        self.torch = torch
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(nstreams)]
        self.slots = [{"arena": None, "graphs": {}} for _ in self.streams]
        self.next_stream = 0
        self.buckets = {}
        self.use_graphs = use_graphs
        self.seen = collections.Counter()  # gradient batch shapes run so far
        self.tomos = {}  # (n, theta, mask_r) -> Tomo plan
        self.arena_growths = 0

    # ---- geometry buckets
    def _bucket(self, geo):
        if geo not in self.buckets:
            from types import SimpleNamespace

            from .bh import BHGradient
            from .rec import Rec
            args = SimpleNamespace(n=geo.n, nz=geo.nz, nobj=geo.nobj, nzobj=geo.nzobj, ndist=geo.ndist,
                                   ntheta=1, energy=geo.energy, z1=np.asarray(geo.z1[:geo.ndist]),
                                   focustodetectordistance=geo.focustodetectordistance,
                                   detector_pixelsize=geo.detector_pixelsize, obj_dtype=geo.obj_dtype,
                                   rho=[1.0, 1.0, 1.0])
            rec = Rec(args)
            self.buckets[geo] = {"rec": rec, "bh": BHGradient(rec)}
        return self.buckets[geo]

    def footprint(self, req):
        """Device bytes a request holds while in flight beyond the arenas: the
        packed batch and output of a propagation, the result copies of a
        gradient, the scalar of a bilinear form."""
        if isinstance(req, TomoRequest):
            # per slice: the stacked copy of the input, the output and the
            # Radon workspace (htamd_tomo_workspace_size: 6 n^2 + ntheta n
            # complex64), with the request's own input counted once more
            nth, n = len(req.theta), req.n
            esz = 8 if req.x.is_complex() else 4
            slab_in = n * n * esz if req.op == "R" else nth * n * esz
            slab_out = nth * n * esz if req.op == "R" else n * n * esz
            return req.nangles * (2 * slab_in + slab_out + 8 * (6 * n * n + nth * n))
        g = req.geometry
        if isinstance(req, OperatorRequest):
            return 2 * req.nangles * g.nz * g.n * 8
        if isinstance(req, BilinearRequest):
            return 1 << 10
        osz = 8 if g.obj_dtype == "complex64" else 4
        return req.nangles * (g.nobj * g.nzobj * osz + g.ndist * 8) + g.ndist * g.n * g.nz * 8

    # ---- arenas
    def _alloc(self, make):
        """make(); if the device is out of memory (results held by clients
        count against it), return the allocator's free blocks to the device and
        try once more."""
        try:
            return make()
        except self.torch.OutOfMemoryError:
            self.torch.cuda.empty_cache()
            return make()

    def _arena(self, si, nbytes):
        slot = self.slots[si]
        if slot["arena"] is None or slot["arena"].numel() < nbytes:
            if slot["graphs"]:
                self.streams[si].synchronize()  # no replay of a graph being dropped is still running
                slot["graphs"].clear()
            slot["arena"] = None  # freed on this stream: reuse by later work on it is ordered
            self.arena_growths += 1
            slot["arena"] = self._alloc(lambda: self.torch.empty(nbytes, dtype=self.torch.uint8,
                                                                 device=self.device))
        return slot["arena"]

    def _views(self, arena, parts):
        """{name: view} of (name, shape, dtype) parts laid out from offset 0,
        each at a 256-byte boundary."""
        out, off = {}, 0
        for name, shape, dt in parts:
            nb = int(np.prod(shape)) * self.torch.empty((), dtype=dt).element_size()
            out[name] = arena[off:off + nb].view(dt).view(shape)
            off += (nb + 255) // 256 * 256
        return out

    @staticmethod
    def _span(parts, element_size):
        return sum((int(np.prod(sh)) * element_size[dt] + 255) // 256 * 256 for _, sh, dt in parts)

    def _staging_parts(self, geo, A, P):
        torch = self.torch
        cdt = torch.complex64 if geo.obj_dtype == "complex64" else torch.float32
        f32, nd = torch.float32, geo.ndist
        del cdt
        return [("pos", (A, nd, 2), f32), ("gpos", (A, nd, 2), f32), ("mag", (A, nd), f32),
                ("tab", (3 * A + 2 * P,), torch.int64), ("scales", (A, 4), f32), ("pidx", (A,), torch.int32),
                ("rho_prb", (P,), f32)]

    def _gradient_arena(self, si, b, geo, A, P):
        torch = self.torch
        parts = self._staging_parts(geo, A, P)
        esz = {torch.complex64: 8, torch.float32: 4, torch.int32: 4, torch.int64: 8}
        span = self._span(parts, esz)
        ws = b["bh"].workspace_bytes(A)
        arena = self._arena(si, span + ws)
        st = self._views(arena, parts)
        st["ws"] = arena[span:span + ws]
        return st

    def _launch(self, b, st, A, P):
        b["bh"].call_tab(st["tab"], A, P, st["pos"], st["mag"], st["pidx"], st["scales"], st["rho_prb"], st["gpos"],
                         st["ws"])

    # ---- batches
    def _run_prop(self, key, batch, s):
        torch = self.torch
        geo, _, op, j = key
        b = self._bucket(geo)
        with torch.cuda.stream(s):
            self._wait_inputs(batch, s)
            for r in batch:
                r.x.record_stream(s)

            def apply():
                x = batch[0].x if len(batch) == 1 else torch.cat([r.x for r in batch])
                return b["rec"].cl_prop.DT(x, j) if op == "DT" else b["rec"].cl_prop.D(x, j)

            y = self._alloc(apply)
            outs, o = [], 0
            for r in batch:
                outs.append(y[o:o + r.nangles])
                o += r.nangles
            ev = torch.cuda.Event()
            ev.record(s)
        return ev, outs

    def _tomo(self, key):
        _, n, theta, mask_r = key[:4]
        if (n, theta, mask_r) not in self.tomos:
            from .tomo import Tomo
            self.tomos[(n, theta, mask_r)] = Tomo(n, 1 << 30, np.asarray(theta), mask_r)
        return self.tomos[(n, theta, mask_r)]

    def _run_tomo(self, key, batch, s):
        torch = self.torch
        op = key[4]
        tomo = self._tomo(key)
        with torch.cuda.stream(s):
            self._wait_inputs(batch, s)
            for r in batch:
                r.x.record_stream(s)
            zdim = 0 if op == "R" else 1

            def apply():
                x = batch[0].x if len(batch) == 1 else torch.cat([r.x for r in batch], dim=zdim)
                return tomo.R(x) if op == "R" else tomo.RT(x)

            y = self._alloc(apply)
            outs, o = [], 0
            for r in batch:  # R output [ntheta, nz, n]: slabs along dim 1; RT output [nz, n, n]
                outs.append(y[:, o:o + r.nangles] if op == "R" else y[o:o + r.nangles])
                o += r.nangles
            ev = torch.cuda.Event()
            ev.record(s)
        return ev, outs

    def _run_hess(self, key, batch, s, si):
        torch = self.torch
        geo = key[0]
        b = self._bucket(geo)
        outs = []
        with torch.cuda.stream(s):
            self._wait_inputs(batch, s)
            mag = b["rec"].eff_demagnifications[0]
            for r in batch:
                for d in (r.x, r.y) + ((r.z,) if r.z is not None else ()):
                    for t in d.values():
                        t.record_stream(s)
                r.data.record_stream(s)
                m = mag.expand(r.nangles, -1).contiguous()
                out = torch.zeros(1, dtype=torch.float64, device=self.device)
                need = b["bh"].hessian_workspace_bytes(r.nangles, r.z is None)
                ws = self._arena(si, need)[:need]
                b["bh"].hessian(r.x, r.y, r.y if r.z is None else r.z, r.data, m, out, accumulate=True,
                                scale=2.0 / r.data_size, workspace=ws)
                outs.append(out)
            ev = torch.cuda.Event()
            ev.record(s)
        return ev, outs

    def run(self, geo, batch):
        torch = self.torch
        si = self.next_stream
        self.next_stream = (si + 1) % len(self.streams)
        s = self.streams[si]
        if not isinstance(geo, Geometry):  # operator and bilinear-form buckets
            if geo[0] == "tomo":
                return self._run_tomo(geo, batch, s)
            return self._run_prop(geo, batch, s) if geo[1] == "prop" else self._run_hess(geo, batch, s, si)
        b = self._bucket(geo)
        A = sum(r.nangles for r in batch)
        P = len(batch)
        slot = self.slots[si]
        with torch.cuda.stream(s):
            self._wait_inputs(batch, s)
            st = self._gradient_arena(si, b, geo, A, P)
            mag = b["rec"].eff_demagnifications[0]
            # the fused passes read each request's projections, data and probe
            # and write its object and probe gradients where they lie, through
            # a table of addresses; only the small per-angle arrays are packed
            o = 0
            tab = [[], [], [], [], []]  # proj, data, gproj, prb, gprb
            outs = []
            cdt = torch.complex64 if geo.obj_dtype == "complex64" else torch.float32
            for p, r in enumerate(batch):
                a = r.nangles
                r.proj = r.proj.to(device=self.device, dtype=cdt).contiguous()
                r.data = r.data.to(device=self.device, dtype=torch.float32).contiguous()
                r.prb = r.prb.to(device=self.device, dtype=torch.complex64).contiguous()
                for t in (r.prb, r.proj, r.pos, r.data):
                    t.record_stream(s)
                gproj = torch.empty_like(r.proj)
                gprb = torch.empty_like(r.prb)
                for t, col in ((r.proj, 0), (r.data, 1), (gproj, 2)):
                    step = t[0].numel() * t.element_size()
                    tab[col].extend(t.data_ptr() + i * step for i in range(a))
                tab[3].append(r.prb.data_ptr())
                tab[4].append(gprb.data_ptr())
                outs.append([gprb, gproj, None])
                st["pos"][o:o + a].copy_(r.pos, non_blocking=True)
                st["mag"][o:o + a].copy_(mag.expand(a, -1), non_blocking=True)
                for c, v in enumerate((2.0 / r.data_size, r.rho[0] ** 2, r.rho[2] ** 2, 0.0)):
                    st["scales"][o:o + a, c].fill_(v)  # device fills: no host copies on the stream
                st["pidx"][o:o + a].fill_(p)
                st["rho_prb"][p].fill_(r.rho[1] ** 2)
                o += a
            # from pinned memory, which the host allocator keeps until the copy has run
            st["tab"].copy_(torch.tensor(sum(tab, []), dtype=torch.int64).pin_memory(), non_blocking=True)
            key = (geo, A, P)
            g = slot["graphs"].get(key) if self.use_graphs else None
            if g is not None:
                g.replay()
            elif self.use_graphs and self.seen[key] >= 1:
                # capture on this stream once its staging work has drained (the
                # torch.cuda.graph context would also synchronize the whole device
                # and empty the allocator's cache); no Python GC frees meanwhile
                s.synchronize()
                g = torch.cuda.CUDAGraph()
                gc_on = gc.isenabled()
                gc.disable()
                try:
                    g.capture_begin(capture_error_mode="thread_local")
                    try:
                        self._launch(b, st, A, P)
                    finally:
                        g.capture_end()
                finally:
                    if gc_on:
                        gc.enable()
                slot["graphs"][key] = g
                g.replay()
            else:
                self._launch(b, st, A, P)
            self.seen[key] += 1
            o = 0
            for p, r in enumerate(batch):
                a = r.nangles
                outs[p][2] = st["gpos"][o:o + a].clone()
                outs[p] = tuple(outs[p])
                o += a
            ev = torch.cuda.Event()
            ev.record(s)
        return ev, outs

    def prepare(self, req):
        """Called in the submitting thread.  Its current stream is where the
        request's inputs were produced and where its results will be used: an
        event recorded there now orders every executor-stream read of the
        inputs after the work the submitter had queued (the executor streams
        do not otherwise synchronise with the submitter's stream)."""
        torch = self.torch
        req.consumer = torch.cuda.current_stream(self.device)
        ev = torch.cuda.Event()
        ev.record(req.consumer)
        req.ready = ev

    @staticmethod
    def _wait_inputs(batch, s):
        """Make stream s wait for each request's inputs (see prepare)."""
        for r in batch:
            ev = getattr(r, "ready", None)
            if ev is not None:
                s.wait_event(ev)

    def done(self, handle):
        return handle[0].query()

    def finish(self, handle, batch):
        """Results were allocated on an executor stream; recording them on the
        consumer's stream keeps the allocator from handing their blocks to a
        later batch while the consumer's kernels may still read them."""
        ev, outs = handle
        ev.synchronize()
        for r, res in zip(batch, outs):
            consumer = getattr(r, "consumer", None)
            if consumer is not None:
                for t in (res if isinstance(res, tuple) else (res,)):
                    t.record_stream(consumer)
            r.future.set_result(res)


class BHServer:
    """Thread-backed request loop: submit() returns a Future of the request's
    result: (gprb [ndist, nz, n], gproj [A, nzobj, nobj], gpos [A, ndist, 2])
    for a GradientRequest, the [B, nz, n] field for an OperatorRequest, the
    request's slab of R / RT output for a TomoRequest, a float64 [1] tensor for
    a BilinearRequest."""

    def __init__(self, hbm_budget=None, max_batch_angles=128, use_graphs=True, executor=None):
        """hbm_budget: device bytes that requests in flight may hold beyond the
        executor's arenas (default a quarter of what is free at start)."""
        if executor is None:
            import torch
            executor = GpuExecutor(use_graphs=use_graphs)
            if hbm_budget is None:
                free, _ = torch.cuda.mem_get_info()
                hbm_budget = free // 4
        self.sched = Scheduler(executor, hbm_budget if hbm_budget is not None else 1 << 40, max_batch_angles)
        self._lock = threading.Condition()
        self._stop = False
        self._thread = threading.Thread(target=self._loop, daemon=True)
        self._thread.start()

    def submit(self, req):
        with self._lock:
            fut = self.sched.submit(req)
            self._lock.notify()
        return fut

    def submit_many(self, reqs):
        """Submit a group atomically: the loop sees all of them at once, so a
        group's batches do not depend on thread timing.  Returns the futures."""
        with self._lock:
            futs = [self.sched.submit(r) for r in reqs]
            self._lock.notify()
        return futs

    def _loop(self):
        while True:
            with self._lock:
                if self._stop and self.sched.idle():
                    return
                moved = self.sched.step()
                if not moved:
                    self._lock.wait(timeout=0.0005)

    def close(self):
        with self._lock:
            self._stop = True
            self._lock.notify()
        self._thread.join()
