"""Continuous-batching request loop for BH operator solves on one GPU.

Three kinds of requests: BH gradients (the main kind, below), propagations
D / DT of batches of fields (packed per geometry, operator and distance into
one launch) and BH bilinear forms (run by the fused bilinear-form kernel, one
launch sequence per request).  Gradient requests (a chunk of angles of some
reconstruction: probe, projected object, positions, measured amplitudes,
geometry) are
  * bucketed by geometry (detector and object grid, distances, physics), since
    one fused launch needs one propagator and one shift plan;
  * admitted while their device footprint (work field, staging, outputs) fits
    the HBM budget, in arrival order within a bucket;
  * packed, as they are admitted, into the bucket's device-resident staging
    arrays (angles of every request contiguous, each request with its own probe
    slot, its own 2/data_size and rho scales), so one fused gradient launch
    (holotomocupy_amd.bh, htamd_bh_gradient) serves the whole batch;
  * launched on one of two streams, so packing and unpacking of one batch
    overlap the kernels of the previous one; the fused launch sequence of a
    batch shape (angles, probes) is captured once as a HIP graph and replayed.
A request's result does not depend on the batch it shares: every per-angle
quantity is computed by the same arithmetic, and probe gradients are summed
per probe slot in angle order (deterministic kernels, no atomics).

The scheduling core (`Scheduler`) is plain Python and is exercised on the CPU
with a fake executor; `GpuExecutor` runs the HIP kernels.
"""
import collections
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

    def _retire(self):
        keep = []
        for h, geo, batch in self.inflight:
            if self.ex.done(h):
                self.ex.finish(h, batch)
                now = time.perf_counter()
                for r in batch:
                    r.t_done = now
                    self.in_use -= r.nbytes
            else:
                keep.append((h, geo, batch))
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
            self.inflight.append((self.ex.run(geo, batch), geo, batch))
        return (len(self.pending), len(self.inflight), self.in_use) != before

    def idle(self):
        return not self.pending and not self.inflight and not any(self.admitted.values())


class GpuExecutor:
    """Runs batches with the fused BH gradient; per-geometry device-resident
    staging and one captured HIP graph per batch shape."""

    def __init__(self, use_graphs=True, nstreams=2):
        import torch
        self.torch = torch
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(nstreams)]
        self.next_stream = 0
        self.buckets = {}
        self.use_graphs = use_graphs

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
            self.buckets[geo] = {"rec": rec, "bh": BHGradient(rec),
                                 "slots": [{"staging": {}, "graphs": {}, "ws": None} for _ in self.streams]}
        return self.buckets[geo]

    def footprint(self, req):
        g = req.geometry
        if isinstance(req, OperatorRequest):
            return 2 * req.nangles * g.nz * g.n * 8  # the packed batch and its output
        if isinstance(req, BilinearRequest):
            nf = 4 if req.z is not None else 3
            osz = 8 if g.obj_dtype == "complex64" else 4
            return req.nangles * (nf * g.ndist * g.n * g.nz * 8 + 3 * g.nobj * g.nzobj * osz) + (1 << 20)
        A, osz = req.nangles, 8 if g.obj_dtype == "complex64" else 4
        field = g.n * g.nz
        plane = g.nobj * g.nzobj
        work = A * g.ndist * field * 8 + 2 * A * plane * osz + A * g.ndist * g.nz * 8
        staging = A * (plane * osz * 2 + g.ndist * field * 4 + g.ndist * 16 + 16) + 2 * g.ndist * field * 8
        return work + staging

    def _staging(self, slot, geo, A, P):
        """Device arrays of a batch of A angles and P probes (grown on demand),
        one set per stream so concurrent batches of a geometry do not share them."""
        torch = self.torch
        st = slot["staging"]
        if st.get("A", 0) < A or st.get("P", 0) < P:
            A2, P2 = max(A, st.get("A", 0)), max(P, st.get("P", 0))
            cdt = torch.complex64 if geo.obj_dtype == "complex64" else torch.float32
            dev = self.device
            st.clear()
            st.update(A=A2, P=P2,
                      proj=torch.empty((A2, geo.nzobj, geo.nobj), dtype=cdt, device=dev),
                      gproj=torch.empty((A2, geo.nzobj, geo.nobj), dtype=cdt, device=dev),
                      pos=torch.empty((A2, geo.ndist, 2), dtype=torch.float32, device=dev),
                      gpos=torch.empty((A2, geo.ndist, 2), dtype=torch.float32, device=dev),
                      mag=torch.empty((A2, geo.ndist), dtype=torch.float32, device=dev),
                      data=torch.empty((A2, geo.ndist, geo.nz, geo.n), dtype=torch.float32, device=dev),
                      scales=torch.empty((A2, 4), dtype=torch.float32, device=dev),
                      pidx=torch.empty((A2,), dtype=torch.int32, device=dev),
                      prb=torch.empty((P2, geo.ndist, geo.nz, geo.n), dtype=torch.complex64, device=dev),
                      gprb=torch.empty((P2, geo.ndist, geo.nz, geo.n), dtype=torch.complex64, device=dev),
                      rho_prb=torch.empty((P2,), dtype=torch.float32, device=dev))
            slot["graphs"].clear()
        return st

    def _workspace(self, b, slot, A):
        need = b["bh"].workspace_bytes(A)
        if slot["ws"] is None or slot["ws"].numel() < need:
            slot["ws"] = self.torch.empty(need, dtype=self.torch.uint8, device=self.device)
            slot["graphs"].clear()
        return slot["ws"]

    def _launch(self, b, slot, A, P):
        st = slot["staging"]
        b["bh"](st["prb"][:P], st["proj"][:A], st["pos"][:A], st["data"][:A], st["mag"][:A], st["gproj"][:A],
                st["gpos"][:A], st["gprb"][:P], accumulate_prb=False, probe_index=st["pidx"][:A],
                scales=st["scales"][:A], rho_prb=st["rho_prb"][:P], workspace=slot["ws"])

    def _run_prop(self, key, batch, s):
        torch = self.torch
        geo, _, op, j = key
        b = self._bucket(geo)
        with torch.cuda.stream(s):
            for r in batch:
                r.x.record_stream(s)
            x = batch[0].x if len(batch) == 1 else torch.cat([r.x for r in batch])
            y = b["rec"].cl_prop.DT(x, j) if op == "DT" else b["rec"].cl_prop.D(x, j)
            outs, o = [], 0
            for r in batch:
                outs.append(y[o:o + r.nangles])
                o += r.nangles
            ev = torch.cuda.Event()
            ev.record(s)
        return ev, outs

    def _run_hess(self, key, batch, s, si):
        torch = self.torch
        geo = key[0]
        b = self._bucket(geo)
        slot = b["slots"][si]  # each stream has its own workspace
        outs = []
        with torch.cuda.stream(s):
            mag = b["rec"].eff_demagnifications[0]
            for r in batch:
                for d in (r.x, r.y) + ((r.z,) if r.z is not None else ()):
                    for t in d.values():
                        t.record_stream(s)
                r.data.record_stream(s)
                m = mag.expand(r.nangles, -1).contiguous()
                out = torch.zeros(1, dtype=torch.float64, device=self.device)
                need = b["bh"].hessian_workspace_bytes(r.nangles, r.z is None)
                if slot.get("hws") is None or slot["hws"].numel() < need:
                    slot["hws"] = torch.empty(need, dtype=torch.uint8, device=self.device)
                b["bh"].hessian(r.x, r.y, r.y if r.z is None else r.z, r.data, m, out, accumulate=True,
                                scale=2.0 / r.data_size, workspace=slot["hws"])
                outs.append(out)
            ev = torch.cuda.Event()
            ev.record(s)
        return ev, outs

    def run(self, geo, batch):
        torch = self.torch
        if not isinstance(geo, Geometry):  # operator and bilinear-form buckets
            si = self.next_stream
            self.next_stream = (si + 1) % len(self.streams)
            s = self.streams[si]
            return self._run_prop(geo, batch, s) if geo[1] == "prop" else self._run_hess(geo, batch, s, si)
        b = self._bucket(geo)
        A = sum(r.nangles for r in batch)
        P = len(batch)
        si = self.next_stream
        s = self.streams[si]
        self.next_stream = (si + 1) % len(self.streams)
        slot = b["slots"][si]
        with torch.cuda.stream(s):
            st = self._staging(slot, geo, A, P)
            self._workspace(b, slot, A)
            mag = b["rec"].eff_demagnifications[0]
            o = 0
            for p, r in enumerate(batch):
                a = r.nangles
                for t in (r.prb, r.proj, r.pos, r.data):
                    t.record_stream(s)
                st["proj"][o:o + a].copy_(r.proj, non_blocking=True)
                st["pos"][o:o + a].copy_(r.pos, non_blocking=True)
                st["data"][o:o + a].copy_(r.data, non_blocking=True)
                st["mag"][o:o + a].copy_(mag.expand(a, -1), non_blocking=True)
                sc = torch.tensor([2.0 / r.data_size, r.rho[0] ** 2, r.rho[2] ** 2, 0.0], dtype=torch.float32)
                st["scales"][o:o + a].copy_(sc.expand(a, -1), non_blocking=True)
                st["pidx"][o:o + a].fill_(p)
                st["prb"][p].copy_(r.prb, non_blocking=True)
                st["rho_prb"][p].fill_(r.rho[1] ** 2)
                o += a
            key = (A, P)
            if self.use_graphs:
                g = slot["graphs"].get(key)
                if g is None:
                    self._launch(b, slot, A, P)  # eager run first (warm-up)
                    g = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(g, stream=s, capture_error_mode="thread_local"):
                        self._launch(b, slot, A, P)
                    slot["graphs"][key] = g
                g.replay()
            else:
                self._launch(b, slot, A, P)
            outs = []
            o = 0
            for p, r in enumerate(batch):
                a = r.nangles
                outs.append((st["gprb"][p].clone(), st["gproj"][o:o + a].clone(), st["gpos"][o:o + a].clone()))
                o += a
            ev = torch.cuda.Event()
            ev.record(s)
        return ev, outs

    def done(self, handle):
        return handle[0].query()

    def finish(self, handle, batch):
        ev, outs = handle
        ev.synchronize()
        for r, res in zip(batch, outs):
            r.future.set_result(res)


class BHServer:
    """Thread-backed request loop: submit() returns a Future of the request's
    result: (gprb [ndist, nz, n], gproj [A, nzobj, nobj], gpos [A, ndist, 2])
    for a GradientRequest, the [B, nz, n] field for an OperatorRequest, a
    float64 [1] tensor for a BilinearRequest."""

    def __init__(self, hbm_budget=None, max_batch_angles=128, use_graphs=True, executor=None):
        if executor is None:
            import torch
            executor = GpuExecutor(use_graphs=use_graphs)
            if hbm_budget is None:
                free, _ = torch.cuda.mem_get_info()
                hbm_budget = int(0.6 * free)
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
