"""BH gradient evaluations/s at a 2048x2048 detector with 4 distances on one MI355X.

One step = one BH gradient evaluation (rec_mpi.Rec.gradients_cascade: object,
probe and position gradients of the data term) for a request of A angles,
detector 2048x2048, 4 distances, object-plane fields 1024x1024 per angle,
inputs resident in HBM.  `value` counts gradient evaluations per second, one
evaluation = one angle across all distances.

Paths: "plain" runs the reference-structured cascade (one kernel per operator
and distance, F3 recomputed per stage, intermediates materialised);
"served" runs the fused multi-pass gradient behind the request loop when
available.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--angles", type=int, default=16, help="angles per request (A)")
    p.add_argument("--n", type=int, default=2048)
    p.add_argument("--nobj", type=int, default=1024)
    p.add_argument("--ndist", type=int, default=4)
    p.add_argument("--path", choices=["auto", "plain", "fused", "served"], default="auto")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-plain", action="store_true", help="skip the plain-path comparison run")
    p.add_argument("--op", choices=["prop", "tomo", "shift", "bh", "serve-mix", "chunking"], default=None,
                   help="operator benchmark instead of the BH gradient: prop = D then DT, 16 fields of "
                        "256^2 at 4 distances; tomo = R and RT at 256^3, 384 angles; shift = coeff, S and "
                        "Sadj at 384 x 256^2; bh = Rec.BH iterations at 512^3, 720 angles, 1024^2; serve-mix = "
                        "Poisson mix of operator, gradient and bilinear-form requests (n 256..2048); chunking = R "
                        "and RT through gpu_batch from host memory at 256^3, 384 angles")
    p.add_argument("--bh-config", choices=["c3", "c4"], default="c3",
                   help="--op bh sizes: c3 = object 512^3, 720 angles, 1024^2 detector; c4 = object 1024^3, "
                        "1800 angles, 2048^2 detector (4 distances both)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="first run a 256 MiB device-to-device copy 3 times (a known-size kernel for "
                        "scaling FETCH_SIZE / WRITE_SIZE counters)")
    return p.parse_args()


def geometry(n, ndist):
    """Physics of the reference's synthetic test (tests/test.py:14-19) at bin 0."""
    import numpy as np
    z1 = np.array([5.110, 5.464, 6.879, 9.817])[:ndist] * 1e-3
    return dict(energy=17.1, detector_pixelsize=1.4760147601476e-6 * 2 * (2048 // n),
                focustodetectordistance=1.217, z1=z1)


def make_inputs(a, seed=1004):
    import numpy as np
    import torch
    from holotomocupy_amd import phantom
    from holotomocupy_amd.rec import Rec
    from types import SimpleNamespace
    rng = np.random.default_rng(seed)
    g = geometry(a.n, a.ndist)
    args = SimpleNamespace(ndist=a.ndist, ntheta=a.angles, nz=a.n, n=a.n, nzobj=a.nobj, nobj=a.nobj,
                           obj_dtype="complex64", rho=[1.0, 0.05, 0.02], **g)
    rec = Rec(args)
    prb = phantom.probe(rng, a.ndist, a.n, a.n)
    proj = phantom.projections(rng, a.angles, a.nobj, a.nobj)
    pos = phantom.positions(rng, a.angles, a.ndist)
    dev = lambda x: torch.from_numpy(x).cuda()  # noqa: E731
    truth = {"prb": dev(prb), "proj": dev(proj), "pos": dev(pos)}
    with torch.no_grad():
        amp = rec.forward(truth).abs()
        data = (amp * (1 + 0.01 * torch.randn_like(amp))).float().contiguous()
    vars_ = {"prb": truth["prb"] * 0.9, "proj": truth["proj"] * 0.8,
             "pos": truth["pos"] + 0.3 * torch.randn_like(truth["pos"])}
    return rec, vars_, data, (prb, proj, pos, args, g)


def plain_step(rec, vars_, grads, data):
    grads["prb"].zero_()
    rec.gradients_cascade(vars_, grads, data)


def fused_step(rec, vars_, grads, data):
    grads["prb"].zero_()
    rec.gradients_fused(vars_, grads, data)


def served_run(a, rec, vars_, data, nreq, warmup):
    """Submit nreq gradient requests of A angles (the request's own probe,
    projections, positions and data) to the serving loop as one burst; returns
    wall time and per-request latencies (ms).  `warmup` identical bursts run
    first, untimed, so plans, staging and the graphs of every batch shape on
    both streams exist before the timed burst."""
    import torch
    from holotomocupy_amd.serve import BHServer, Geometry, GradientRequest
    g = geometry(a.n, a.ndist)
    geo = Geometry(a.n, a.n, a.nobj, a.nobj, a.ndist, energy=g["energy"], z1=tuple(float(z) for z in g["z1"]),
                   focustodetectordistance=g["focustodetectordistance"], detector_pixelsize=g["detector_pixelsize"])
    srv = BHServer(max_batch_angles=max(a.angles, 64))
    mk = lambda: GradientRequest(geo, vars_["prb"], vars_["proj"], vars_["pos"], data,  # noqa: E731
                                 data_size=rec.data_size, rho=(1.0, 0.05, 0.02))
    for _ in range(max(2, warmup)):  # a burst is submitted atomically, so its batches repeat
        for f in srv.submit_many([mk() for _ in range(nreq)]):
            f.result()
    torch.cuda.synchronize()
    reqs = [mk() for _ in range(nreq)]
    t0 = time.perf_counter()
    futs = srv.submit_many(reqs)
    for f in futs:
        f.result()
    wall = time.perf_counter() - t0
    srv.close()
    lat = sorted((r.t_done - r.t_submit) * 1e3 for r in reqs)
    return wall, lat


def time_steps(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for s, e in evs:
        s.record()
        fn()
        e.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = sorted(s.elapsed_time(e) for s, e in evs)
    return wall, per


def roofline_fresnel_cols(rec, a, reps=20):
    """Dominant kernel of the plain path: the Fresnel column pass on a batch of
    A detector fields.  Algorithmic bytes = read + write of each field
    (2 * 8 * nz * n per field).  Timed live with events on the launch stream."""
    import torch
    from holotomocupy_amd import _lib
    B = a.angles
    buf = torch.randn(B, a.n, a.n, dtype=torch.complex64, device="cuda")
    st = torch.cuda.current_stream()
    args = (rec.cl_prop._plan, _lib.ptr(buf), a.n * a.n, _lib.ptr(buf), a.n * a.n, B, 0, 0, 1,
            _lib.stream_ptr(st))
    for _ in range(3):
        _lib.call("htamd_prop_pass", *args)
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record(st)
    for _ in range(reps):
        _lib.call("htamd_prop_pass", *args)
    e.record(st)
    torch.cuda.synchronize()
    us = s.elapsed_time(e) * 1e3 / reps
    nbytes = 2 * 8 * a.n * a.n * B
    achieved = nbytes / (us * 1e-6) / 1e9
    del buf
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "fresnel_cols_k (Fresnel column pass, %d fields of %dx%d)" % (B, a.n, a.n),
            "bytes_per_launch": nbytes, "avg_us": round(us, 2)}


STAGES = ["prefilter", "rows_fwd", "cols", "rows_adj", "sadj", "prefilter_grad", "reductions"]

# HBM bytes per launch of the fused passes at the default workload (2048^2
# detector, 4 distances, 1024^2 object plane, 16 angles), from rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE runs scaled by what the counters read for a 256 MiB
# device-to-device copy (reads x2.00, writes x1.00); the commands and per-kernel
# figures are in profiles/bh_gradient_fused_2048_d4_a16_traffic.md.
TRAFFIC_MIB = {"bh_fwd_rows_k": 1503.4 + 2048.0, "bh_cols_k": 3149.4 + 3399.1, "bh_adj_rows_k": 4037.8 + 2177.6}
TRAFFIC_WORKLOAD = (2048, 4, 1024, 16)


def roofline_fused(a, rec, vars_, grads, data, steps=3):
    """Per-pass durations of the fused gradient, measured with HIP events
    recorded by the library around each pass on the launch stream, over
    `steps` gradient evaluations of the same workload.  The dominant pass is
    reported against its compulsory HBM bytes:
      rows_fwd: A nd F + A Fo + nd F      (write work field; read c, probe)
      cols:     2.5 A nd F                (read/write work field, read sqrt(data))
      rows_adj: 2 A nd F + A Fo + 2 nd F  (read/write work field; read c, probe; write probe gradient)
    with F = 8 nz n, Fo = 8 nzobj nobj bytes."""
    import ctypes
    import torch
    from holotomocupy_amd import _lib
    fused_step(rec, vars_, grads, data)  # plan + workspace
    bh = rec._bh
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(2 * len(STAGES))]
    for e in evs:  # torch creates the HIP event on first record
        e.record()
    torch.cuda.synchronize()
    arr = (ctypes.c_void_p * len(evs))(*[ctypes.c_void_p(e.cuda_event) for e in evs])
    if not all(x for x in arr):
        raise RuntimeError("stage events were not created")
    _lib.call("htamd_bh_set_stage_events", bh._plan, arr, len(evs))
    tot = [0.0] * len(STAGES)
    try:
        for _ in range(steps):
            fused_step(rec, vars_, grads, data)
            torch.cuda.synchronize()
            for k in range(len(STAGES)):
                tot[k] += evs[2 * k].elapsed_time(evs[2 * k + 1]) / steps
    finally:
        _lib.call("htamd_bh_set_stage_events", bh._plan, None, 0)
    A, nd = a.angles, a.ndist
    F, Fo = 8 * a.n * a.n, 8 * a.nobj * a.nobj
    model = {"rows_fwd": A * nd * F + A * Fo + nd * F, "cols": 2.5 * A * nd * F,
             "rows_adj": 2 * A * nd * F + A * Fo + 2 * nd * F}
    k = max(model, key=lambda s: tot[STAGES.index(s)])
    us = tot[STAGES.index(k)] * 1e3
    achieved = model[k] / (us * 1e-6) / 1e9
    kname = {"rows_fwd": "bh_fwd_rows_k", "cols": "bh_cols_k", "rows_adj": "bh_adj_rows_k"}[k]
    traffic = None
    if (a.n, a.ndist, a.nobj, a.angles) == TRAFFIC_WORKLOAD:
        traffic = int(TRAFFIC_MIB[kname] * 2 ** 20)
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": kname,
            "bytes_per_launch": int(model[k]), "avg_us": round(us, 2),
            "pass_ms": {s: round(t, 3) for s, t in zip(STAGES, tot)}}


def cpu_baseline(a, host, vars_, data, runs=5):
    """Oracle (NumPy/SciPy port of the reference cascade) on one angle of the
    same workload: one BH gradient evaluation at the timed variables and data
    (angle 0), in FP32 (complex64 transforms and sums, the reference's
    precision) and in FP64, each timed as the median of `runs` runs after one
    warm-up (the FP64 warm-up run also gives the reference gradient).  Returns
    the baseline record and (inputs, float32 data, FP64 gradients) for the
    parity check of the GPU path on the same sample."""
    import numpy as np
    from oracle import cascade as oc
    from oracle import propagation as op
    from oracle import shift as osh
    from holotomocupy_amd.rec import physics
    _, _, _, args, g = host
    cores, model = _host_cpu()
    op.WORKERS = cores
    osh.WORKERS = cores
    ph = physics(g["energy"], g["z1"], g["focustodetectordistance"], g["detector_pixelsize"])
    mag = (1 / ph.norm_magnifications)[None, :]
    prb = vars_["prb"].cpu().numpy()
    proj = vars_["proj"][:1].cpu().numpy()
    pos = vars_["pos"][:1].cpu().numpy()
    d = data[:1].cpu().numpy()
    times, ref = {}, None
    for single in (False, True):
        cd = np.complex64 if single else np.complex128
        prop = op.Propagation(a.n, a.n, 1, a.ndist, ph.wavelength, ph.voxelsize, ph.distance, dtype=cd)
        cas = oc.Cascade(prop, osh.Shift(a.n, a.nobj, a.n, a.nobj, single=single), a.ndist, a.n, a.n, a.nobj,
                         a.nobj, a.angles * a.ndist * a.n * a.n, mag, rho=tuple(args.rho), single=single)
        x = [prb.astype(cd), proj.astype(cd), pos.astype(np.float64)]
        dd = d.astype(np.float32 if single else np.float64)
        ts = []
        for k in range(runs + 1):
            t0 = time.perf_counter()
            out = cas.gradients(*x, dd)
            if k:
                ts.append(time.perf_counter() - t0)
            elif not single:
                ref = out
            print(f"cpu_baseline: {'FP32' if single else 'FP64'} run {k + 1}/{runs + 1} done", file=sys.stderr,
                  flush=True)
        times["fp32" if single else "fp64"] = float(np.median(ts))
    t32, t64 = times["fp32"], times["fp64"]
    base = {"value": round(1.0 / t32, 5), "unit": "evals/s", "cores": cores, "cpu_model": model, "kind": "port",
            "precision": "fp32", "fp64": {"value": round(1.0 / t64, 5), "unit": "evals/s", "median_s": round(t64, 3)},
            "sample": f"1 angle x {a.ndist} distances at {a.n}^2 / object {a.nobj}^2, NumPy/SciPy oracle cascade "
                      f"(scipy.fft workers={cores}), median of {runs} runs after 1 warm-up: FP32 {t32:.2f} s, "
                      f"FP64 {t64:.2f} s"}
    return base, (prb, proj, pos, d, ref)


def parity_vs_fp64(rec, sample):
    """Relative L2 error against the FP64 oracle on the cpu_baseline sample
    (same inputs, same float32 data) of the fused GPU gradient and, for
    comparison, of the per-operator FP32 GPU path."""
    import numpy as np
    import torch
    prb, proj, pos, d, ref = sample
    dev = lambda v: torch.from_numpy(np.ascontiguousarray(v)).cuda()  # noqa: E731
    x = {"prb": dev(prb), "proj": dev(proj), "pos": dev(pos)}
    out = {}
    for path in ("fused", "cascade"):
        g = {"prb": torch.zeros_like(x["prb"]), "proj": torch.empty_like(x["proj"]),
             "pos": torch.empty_like(x["pos"])}
        getattr(rec, f"gradients_{path}")(x, g, dev(d))
        torch.cuda.synchronize()
        errs = {}
        for k, r in zip(("prb", "proj", "pos"), ref):
            got = g[k].cpu().numpy().astype(np.complex128)
            errs[k] = float(np.linalg.norm(got - r) / np.linalg.norm(r))
        out[path if path == "fused" else "per_operator"] = errs
    out["tolerance"] = "1e-5 (prb, proj), 1e-4 (pos)"
    return out


def measured_copy_gbs(nbytes=1 << 30, reps=10):
    """Streaming device-to-device copy on this GPU: (bytes read + bytes
    written) / time, mean over `reps` copies of nbytes after two warm-ups."""
    import torch
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return round(2 * nbytes / (ms * 1e-3) / 1e9, 1)


def _host_cpu():
    """(threads used by the oracle, CPU model) of this host: lscpu's model
    name with its vendor family / model numbers (a virtual machine may report
    only a generic name)."""
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cores = max(1, min(cores, os.cpu_count() or 1))
    fields = {}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                fields[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    if "Model name" not in fields:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        fields["Model name"] = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    model = fields.get("Model name", "unknown")
    ids = [f"{k.lower()} {fields[k]}" for k in ("CPU family", "Model", "Stepping") if k in fields]
    if ids:
        model += " (" + ", ".join(ids) + ")"
    if "Socket(s)" in fields and "Core(s) per socket" in fields:
        model += f", {fields['Socket(s)']} socket(s) x {fields['Core(s) per socket']} cores"
    return cores, model


def _op_cpu_baseline(fn, evals, unit, sample):
    """Time fn() (the oracle on a bounded sample worth `evals` units) once
    after one warm-up call."""
    from oracle import propagation as op
    from oracle import shift as osh
    cores, model = _host_cpu()
    op.WORKERS = osh.WORKERS = cores
    fn()
    t0 = time.perf_counter()
    fn()
    dt = time.perf_counter() - t0
    return {"value": round(evals / dt, 5), "unit": unit, "cores": cores, "cpu_model": model, "kind": "port",
            "sample": f"{sample}; FP64 NumPy/SciPy oracle, one run of {dt:.2f} s"}


def _timed(fn, steps, warmup):
    """Mean ms per call of fn over `steps` calls after `warmup`, with events on
    the current stream."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(steps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / steps


def _op_line(metric, value, unit, ms, nbytes, workload, a, parts):
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"metric": metric, "value": round(value, 3), "unit": unit, "n_gpus": 1, "steps": a.steps,
            "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "c64", "data": "synthetic (seeded)",
            "config": {"workload": workload},
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                         "bytes_per_step": int(nbytes)},
            "parts_ms": {k: round(v, 4) for k, v in parts.items()}}


def op_prop(a):
    """D then DT on 16 complex64 fields of 256^2 at 4 distances (Fresnel numbers
    1e-3, 1.2e-3, 1.5e-3, 2e-3 with voxelsize = wavelength = 1).  Bytes per
    field and operator: 4F (row pass and column pass, F = 8 nz n)."""
    import numpy as np
    import torch
    from holotomocupy_amd import Propagation
    n, B, nd = 256, 16, 4
    prop = Propagation(n, n, B, nd, 1.0, 1.0, 1.0 / np.array([1e-3, 1.2e-3, 1.5e-3, 2e-3]))
    g = torch.Generator(device="cuda").manual_seed(1001)
    x = torch.randn((B, n, n), dtype=torch.complex64, device="cuda", generator=g)
    y = torch.empty_like(x)

    def step():
        for j in range(nd):
            prop.apply_into(x, y, j, False)
            prop.apply_into(y, y, j, True)

    ms_eager = _timed(step, a.steps, a.warmup)
    ref = y.clone()
    # the step's 16 launches (eight calls of two passes each) are launch-bound at
    # 256^2; captured once in a HIP graph they replay as one submission
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step()
    for _ in range(3):  # every replay equals the eager step, with eager work in between
        y.zero_()
        graph.replay()
        torch.cuda.synchronize()
        if not torch.equal(y, ref):
            raise RuntimeError("graph replay of the D/DT step differs from the eager step")
        step()
    ms = _timed(graph.replay, a.steps, a.warmup)
    F = 8 * n * n
    nbytes = nd * 2 * B * 4 * F
    line = _op_line("D+DT field evals/s (16 x 256^2, 4 distances)", nd * 2 * B / (ms * 1e-3), "evals/s", ms,
                    nbytes, "Fresnel D then DT, 16 fields 256^2, 4 distances (one HIP graph per step)", a,
                    {"step": ms, "step_eager": ms_eager})
    line["eager_value"] = round(nd * 2 * B / (ms_eager * 1e-3), 3)
    if not a.no_cpu_baseline:
        from oracle import propagation as op
        o = op.Propagation(n, n, B, nd, 1.0, 1.0, 1.0 / np.array([1e-3, 1.2e-3, 1.5e-3, 2e-3]))
        xs = x.cpu().numpy().astype(np.complex128)

        def oracle_step():
            for j in range(nd):
                o.DT(o.D(xs, j), j)

        line["cpu_baseline"] = _op_cpu_baseline(oracle_step, nd * 2 * B, "evals/s",
                                                "the 16 fields, D then DT at 4 distances")
    return line


def op_tomo(a):
    """R and RT at 256^3 with 384 angles; bytes per slice 8 (13 n^2 + ntheta n)
    for each operator."""
    import numpy as np
    import torch
    from holotomocupy_amd import Tomo
    n, nth = 256, 384
    tomo = Tomo(n, n, np.arange(nth) * np.pi / nth, 0)
    g = torch.Generator(device="cuda").manual_seed(1002)
    u = torch.randn((n, n, n), dtype=torch.complex64, device="cuda", generator=g)
    d = torch.randn((nth, n, n), dtype=torch.complex64, device="cuda", generator=g)
    ms_r = _timed(lambda: tomo.R(u), a.steps, a.warmup)
    ms_rt = _timed(lambda: tomo.RT(d), a.steps, a.warmup)
    nbytes = 2 * n * 8 * (13 * n * n + nth * n)
    ms = ms_r + ms_rt
    line = _op_line("R+RT evals/s (256^3, 384 angles)", 1.0 / (ms * 1e-3), "evals/s", ms, nbytes,
                    "USFFT Radon R and RT, 256 slices of 256^2, 384 angles", a, {"R": ms_r, "RT": ms_rt})
    if not a.no_cpu_baseline:
        from oracle import tomo as ot
        ns = 4
        o = ot.Tomo(n, ns, np.arange(nth) * np.pi / nth, 0)
        us = u[:ns].cpu().numpy().astype(np.complex128)
        ds = d[:, :ns].cpu().numpy().astype(np.complex128)
        line["cpu_baseline"] = _op_cpu_baseline(lambda: (o.R(us), o.RT(ds)), ns / n, "evals/s",
                                                f"{ns} of the {n} slices (R and RT), scaled to the volume")
    return line


def op_shift(a):
    """coeff, S and Sadj on 384 complex64 planes of 256^2 (m = 1, shifts
    U[-5, 5]); bytes: coeff 2 Fo, S Fo + F, Sadj F + Fo per plane."""
    import numpy as np
    import torch
    from holotomocupy_amd import Shift
    n, B = 256, 384
    sh = Shift(n, n, n, n, "complex64")
    rng = np.random.default_rng(1002)
    g = torch.Generator(device="cuda").manual_seed(1002)
    psi = torch.randn((B, n, n), dtype=torch.complex64, device="cuda", generator=g)
    r = torch.from_numpy(rng.uniform(-5, 5, (B, 2)).astype(np.float32)).cuda()
    m = torch.ones(B, dtype=torch.float32, device="cuda")
    c = sh.coeff(psi)
    ms_c = _timed(lambda: sh.coeff(psi), a.steps, a.warmup)
    ms_s = _timed(lambda: sh.S(c, r, m), a.steps, a.warmup)
    ms_a = _timed(lambda: sh.Sadj(psi, r, m), a.steps, a.warmup)
    F = 8 * n * n
    nbytes = B * (2 * F + 2 * F + 2 * F)
    ms = ms_c + ms_s + ms_a
    line = _op_line("coeff+S+Sadj plane evals/s (384 x 256^2)", B / (ms * 1e-3), "evals/s", ms, nbytes,
                    "B-spline prefilter, shift and adjoint shift, 384 planes of 256^2", a,
                    {"coeff": ms_c, "S": ms_s, "Sadj": ms_a})
    if not a.no_cpu_baseline:
        from oracle import shift as osh
        ns = 128
        o = osh.Shift(n, n, n, n, "complex64")
        ps = psi[:ns].cpu().numpy().astype(np.complex128)
        rs, ms_ = r[:ns].cpu().numpy(), m[:ns].cpu().numpy()

        def oracle_step():
            cs = o.coeff(ps)
            o.S(cs, rs, ms_)
            o.Sadj(ps, rs, ms_)

        line["cpu_baseline"] = _op_cpu_baseline(oracle_step, ns, "evals/s", f"{ns} of the {B} planes")
    return line


def op_chunking(a):
    """Rec.fwd_tomo (R) and Rec.gF4 (RT) over a host-resident 256^3 object and
    384-angle sinogram through Chunking.gpu_batch (chunks of 16 slices, pinned
    double-buffered staging, copies and compute on separate streams ordered
    by events), beside the same operators on device-resident arrays.  Bytes:
    the host <-> device traffic of one pass (object in + sinogram out)."""
    import numpy as np
    import torch
    from types import SimpleNamespace
    from holotomocupy_amd.rec import Rec
    n, nth = 256, 384
    theta = np.arange(nth) * np.pi / nth
    args = SimpleNamespace(ndist=1, ntheta=nth, nz=n, n=n, nzobj=n, nobj=n, obj_dtype="complex64",
                           rho=[1.0, 1.0, 1.0], nchunk=16, theta=theta, mask_r=0, **geometry(n, 1))
    rec = Rec(args)
    rng = np.random.default_rng(1002)
    u = (rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))).astype(np.complex64)
    d = (rng.standard_normal((nth, n, n)) + 1j * rng.standard_normal((nth, n, n))).astype(np.complex64)
    out_h, gu_h = np.empty((nth, n, n), np.complex64), np.empty((n, n, n), np.complex64)
    u_d, d_d = torch.from_numpy(u).cuda(), torch.from_numpy(d).cuda()
    out_d, gu_d = torch.empty((nth, n, n), dtype=torch.complex64, device="cuda"), torch.empty_like(u_d)

    def wall(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3

    reps = max(3, a.steps // 2)
    parts = {"R_host": wall(lambda: rec.fwd_tomo(u, out_h), reps),
             "RT_host": wall(lambda: rec.gF4(gu_h, d), reps),
             "R_device": wall(lambda: rec.fwd_tomo(u_d, out_d), reps),
             "RT_device": wall(lambda: rec.gF4(gu_d, d_d), reps)}
    ms = parts["R_host"] + parts["RT_host"]
    nbytes = 2 * (u.nbytes + d.nbytes)  # each operator moves the object and the sinogram once
    line = _op_line("R+RT evals/s through gpu_batch from host memory (256^3, 384 angles)", 1e3 / ms, "evals/s", ms,
                    nbytes, "Rec.fwd_tomo and Rec.gF4 on host-resident arrays, chunks of 16 slices", a, parts)
    line["roofline"] = {"bound": "pcie", "host_device_gbs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                        "bytes_per_step": int(nbytes), "device_resident_ms": round(parts["R_device"]
                                                                                   + parts["RT_device"], 3)}
    return line


def _bilinear_roofline(a, rec, n, nobj, nd, A=16):
    """The fused bilinear-form passes on one chunk of A angles at the loop's
    point, in the three forms the loop calls (z = y, general, and the beta
    step's pair), timed with events.  Model bytes per angle, with NF the
    propagated fields (3, 4, 5): the prefilters of the 2 or 3 projections
    (read and write each, 2 F_o apiece), the row pass reading those planes once
    and writing NF fields per distance, the column pass reading them back with
    the data (F/2 per distance), and the NF - 1 probes (per distance) read once
    per launch: 2 NF nd F + nd F / 2 + 3 P F_o + (NF - 1) nd F / A, P the
    projections (2 or 3)."""
    import torch
    F, Fo = 8 * n * n, 8 * nobj * nobj
    gen = torch.Generator(device="cuda").manual_seed(7)
    x = {"prb": rec.vars["prb"], "proj": rec.vars["proj"][:A], "pos": rec.vars["pos"][:A]}
    d = rec.data[:A]

    def direction(sc):
        return {"prb": sc * 0.1 * rec.vars["prb"],
                "proj": sc * 0.05 * torch.randn(x["proj"].shape, dtype=x["proj"].dtype, device="cuda", generator=gen),
                "pos": sc * 0.3 * torch.randn(x["pos"].shape, device="cuda", generator=gen)}

    y, z = direction(1.0), direction(0.7)
    out = {}
    for name, nf, nproj, fn in (("same", 3, 2, lambda: rec.hessian_fused(x, y, y, d)),
                                ("general", 4, 3, lambda: rec.hessian_fused(x, y, z, d)),
                                ("pair", 5, 3, lambda: rec.hessian_pair_fused(x, y, z, d))):
        ms = _timed(fn, max(3, a.steps), 1)
        nbytes = A * (2 * nf * nd * F + nd * F / 2 + 3 * nproj * Fo) + (nf - 1) * nd * F
        out[name] = {"ms": round(ms, 3), "bytes": int(nbytes), "achieved_gbs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                     "frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    out["angles"] = A
    return out


def op_bh(a):
    """Rec.BH iterations on one GPU (16-angle launches, 16-slice tomography
    chunks, Laplacian regulariser on): `--bh-config c3` is object 512^3, 720
    angles, detector 1024^2, 4 distances; `c4` is one rank's share of the
    largest case, object 1024^3, 1800 angles, detector 2048^2, 4 distances
    (121 GB of data, every array resident in HBM).  One warm-up iteration (it
    has no beta step), then `steps` timed iterations; the phases are timed with
    events on the launch stream.  Bytes: the gradient's model bytes per angle
    times the angles, for the roofline of the whole iteration."""
    import numpy as np
    import torch
    from types import SimpleNamespace
    from holotomocupy_amd import phantom
    from holotomocupy_amd.rec import Rec
    n, nobj, nth, nd = (2048, 1024, 1800, 4) if a.bh_config == "c4" else (1024, 512, 720, 4)
    say = lambda m: print(f"[bh {a.bh_config}] {m}", file=sys.stderr, flush=True)  # noqa: E731
    g = geometry(n, nd)
    theta = np.arange(nth) * np.pi / nth
    args = SimpleNamespace(ndist=nd, ntheta=nth, nz=n, n=n, nzobj=nobj, nobj=nobj, obj_dtype="complex64",
                           rho=[1.0, 0.05, 0.02], theta=theta, mask_r=0, niter=1 + a.steps, start_iter=0,
                           err_step=-1, vis_step=-1, nchunk=16, lam_prbfit=0.0, lam_laplacian=1e-3, **g)
    rec = Rec(args)
    rng = np.random.default_rng(1003)
    gen = torch.Generator(device="cuda").manual_seed(1003)
    zz, yy, xx = [t.float() - nobj / 2 for t in torch.meshgrid(*(torch.arange(nobj, device="cuda"),) * 3,
                                                                  indexing="ij")]
    delta = torch.zeros((nobj, nobj, nobj), device="cuda")
    for _ in range(30):
        c = (torch.rand(3, generator=gen, device="cuda") - 0.5) * 0.7 * nobj
        ax = (0.05 + 0.2 * torch.rand(3, generator=gen, device="cuda")) * nobj
        inside = ((zz - c[0]) / ax[0]) ** 2 + ((yy - c[1]) / ax[1]) ** 2 + ((xx - c[2]) / ax[2]) ** 2 <= 1
        delta += inside.float() * torch.rand(1, generator=gen, device="cuda")
    del zz, yy, xx
    obj = torch.complex(-delta, delta / 100)
    del delta
    say("phantom made")
    proj = torch.empty((nth, nobj, nobj), dtype=torch.complex64, device="cuda")
    rec.fwd_tomo(obj / float(rec.norm_const), out=proj)
    obj *= 1.0 / float(proj.abs().max())  # phase of order one radian
    del proj
    say("phantom scaled")
    prb = torch.from_numpy(phantom.probe(rng, nd, n, n)).cuda()
    pos = torch.from_numpy(phantom.positions(rng, nth, nd)).cuda()
    rec.gen_sqrt_data({"obj": obj, "prb": prb, "pos": pos})
    say(f"data made; {torch.cuda.memory_allocated() / 2**30:.1f} GiB allocated")
    rec.vars["obj"].copy_(obj * 0.8)
    rec.vars["prb"].copy_(prb * 0.9)
    rec.vars["pos"].copy_(pos + 0.3 * torch.randn(pos.shape, generator=gen, device="cuda"))
    del obj
    rec.time_phases = True
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rec.BH()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    say(f"{1 + a.steps} iterations in {wall:.2f} s; peak {torch.cuda.max_memory_allocated() / 2**30:.1f} GiB")
    timed = rec.phase_ms[1:]
    parts = {k: sum(p[k] for p in timed) / len(timed) for k in timed[0]}
    ms = sum(parts.values())
    F, Fo = 8 * n * n, 8 * nobj * nobj
    nbytes = nth * (nd * (6.5 * F + 2 * Fo) + 9 * Fo)
    shape = f"object {nobj}^3, {nth} angles, {n}^2 detector, {nd} distances"
    line = _op_line(f"BH iterations/s ({shape})", 1e3 / ms, "iterations/s", ms, nbytes,
                    f"Rec.BH: gradient, R of the gradient, three bilinear forms, updates; {shape}, Laplacian on",
                    a, parts)
    line["warmup"] = 1  # the first iteration (no beta step) is the warm-up
    line["bilinear_roofline"] = _bilinear_roofline(a, rec, n, nobj, nd)
    line["peak_hbm_gib"] = round(torch.cuda.max_memory_allocated() / 2**30, 1)
    line["wall_s_all_iterations"] = round(wall, 3)
    line["alphas"] = [float(x) for x in rec.alphas]
    return line


def op_serve_mix(a, nreq=200, rate=200.0, seed=1005):
    """Serving mix: Poisson arrivals at `rate` requests/s of operator (D or DT of
    A fields at one distance, R or RT of an 8-slice slab at A angles), BH
    gradient and BH bilinear-form requests
    (weights 0.4 / 0.4 / 0.2), detector and object grid n in {256, 512, 1024, 2048}
    (weights 0.4 / 0.3 / 0.2 / 0.1, object n x n), 1-4 distances, 16-128
    angles.  Served: BHServer (admission by HBM budget, batches per bucket, HIP
    graphs for gradient batches).  Plain: the same requests one at a time through
    the per-operator path (gradients and bilinear forms over 16-angle chunks)."""
    import threading
    import numpy as np
    import torch
    from holotomocupy_amd import phantom
    from holotomocupy_amd.rec import Rec
    from holotomocupy_amd.serve import (BHServer, BilinearRequest, Geometry, GradientRequest, OperatorRequest,
                                        TomoRequest)
    from holotomocupy_amd.tomo import Tomo
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    sizes, kinds = [256, 512, 1024, 2048], ["op", "grad", "hess"]
    specs = []
    for _ in range(nreq):
        n = int(rng.choice(sizes, p=[0.4, 0.3, 0.2, 0.1]))
        specs.append((str(rng.choice(kinds, p=[0.4, 0.4, 0.2])), n, int(rng.integers(1, 5)),
                      int(rng.integers(16, 129))))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    geos, probes = {}, {}

    def geo(n, nd):
        if (n, nd) not in geos:
            g = geometry(n, nd)
            geos[(n, nd)] = Geometry(n, n, n, n, nd, energy=g["energy"], z1=tuple(float(z) for z in g["z1"]),
                                     focustodetectordistance=g["focustodetectordistance"],
                                     detector_pixelsize=g["detector_pixelsize"])
            probes[(n, nd)] = torch.from_numpy(phantom.probe(rng, nd, n, n)).cuda()
        return geos[(n, nd)]

    pools = {}

    def pool(name, n, count, dtype, scale, shift=0.0):
        """Leading `count` elements of a seeded per-size buffer: requests of one
        grid share their (read-only) inputs, so 200 requests fit in HBM."""
        if (name, n) not in pools:
            numel = 128 * (4 if name == "data" else 1) * n * n
            pools[(name, n)] = shift + scale * torch.randn(numel, dtype=dtype, device="cuda", generator=gen)
        return pools[(name, n)][:count]

    def make(kind, n, nd, A):
        G = geo(n, nd)
        if kind == "op":
            op = str(rng.choice(["D", "DT", "R", "RT"]))
            if op in ("R", "RT"):  # slab of 8 slices, A uniform angles on [0, pi), no mask
                theta = tuple(float(t) for t in np.arange(A) * np.pi / A)
                shape = (8, n, n) if op == "R" else (A, 8, n)
                x = pool("field", n, int(np.prod(shape)), torch.complex64, 1.0).view(shape)
                return TomoRequest(n, theta, 0.0, op, x)
            x = pool("field", n, A * n * n, torch.complex64, 1.0).view(A, n, n)
            return OperatorRequest(G, op, int(rng.integers(0, nd)), x)
        proj = pool("proj", n, A * n * n, torch.complex64, 0.3).view(A, n, n)
        pos = 2 * torch.rand((A, nd, 2), device="cuda", generator=gen) - 1
        data = pool("data", n, A * nd * n * n, torch.float32, 0.25, 1.0).view(A, nd, n, n)
        if kind == "grad":
            return GradientRequest(G, probes[(n, nd)], proj, pos, data, data_size=float(A * nd * n * n),
                                   rho=(1.0, 0.05, 0.02))
        x = {"prb": probes[(n, nd)], "proj": proj, "pos": pos}
        y = {"prb": 0.1 * probes[(n, nd)], "proj": pool("dproj", n, A * n * n, torch.complex64, 0.05).view(A, n, n),
             "pos": 0.3 * torch.randn_like(pos)}
        return BilinearRequest(G, x, y, None, data, float(A * nd * n * n))

    reqs = [make(*sp) for sp in specs]
    torch.cuda.synchronize()
    say = lambda m: print(f"[serve-mix] {m}", file=sys.stderr, flush=True)  # noqa: E731
    say(f"{nreq} requests made")
    clone = lambda r: (OperatorRequest(r.geometry, r.op, r.j, r.x) if isinstance(r, OperatorRequest) else  # noqa: E731
                       TomoRequest(r.n, r.theta, r.mask_r, r.op, r.x) if isinstance(r, TomoRequest) else
                       GradientRequest(r.geometry, r.prb, r.proj, r.pos, r.data, r.data_size, r.rho)
                       if isinstance(r, GradientRequest) else
                       BilinearRequest(r.geometry, r.x, r.y, r.z, r.data, r.data_size))
    srv = BHServer(max_batch_angles=128)

    def run_pass(paced):
        """Submit clones of the requests (at their arrival times if paced) and
        wait for all; results are dropped as they complete, as a consuming
        client would.  Returns (latencies in ms, wall seconds)."""
        timed = [clone(r) for r in reqs]
        lock, finished, errors = threading.Lock(), threading.Event(), []
        left = [len(timed)]

        def on_done(f, r):
            r.t_cb = time.perf_counter()
            if f.exception() is not None:
                errors.append(f.exception())
            r.future = None
            with lock:
                left[0] -= 1
                if left[0] == 0:
                    finished.set()

        t0 = time.perf_counter()

        def feeder():
            for r, t in zip(timed, arrivals):
                d = t0 + t - time.perf_counter()
                if paced and d > 0:
                    time.sleep(d)
                srv.submit(r).add_done_callback(lambda f, r=r: on_done(f, r))

        th = threading.Thread(target=feeder)
        th.start()
        th.join()
        finished.wait()
        wall = time.perf_counter() - t0
        if errors:
            raise errors[0]
        return sorted((r.t_cb - r.t_submit) * 1e3 for r in timed), wall

    try:
        run_pass(True)  # rehearsal with the same arrivals: plans, staging, graphs of the batch shapes
        say("rehearsal done")
        run_pass(False)
        torch.cuda.synchronize()
        say("capacity rehearsal done")
        lat, wall_s = run_pass(True)
        _, cap_s = run_pass(False)
        say(f"served: offered {wall_s:.3f} s, capacity {cap_s:.3f} s, arena growths {srv.sched.ex.arena_growths}")
    finally:
        srv.close()
    # plain: one request at a time, per-operator kernels, 16-angle chunks
    recs = {}

    def rec_for(G):
        if G not in recs:
            g = geometry(G.n, G.ndist)
            from types import SimpleNamespace
            recs[G] = Rec(SimpleNamespace(ndist=G.ndist, ntheta=16, nz=G.n, n=G.n, nzobj=G.n, nobj=G.n,
                                          obj_dtype="complex64", rho=[1.0, 0.05, 0.02], nchunk=16, **g))
        return recs[G]

    tomos = {}

    def plain(r):
        if isinstance(r, TomoRequest):
            if (r.n, r.theta) not in tomos:
                tomos[(r.n, r.theta)] = Tomo(r.n, 1 << 30, np.asarray(r.theta), r.mask_r)
            t = tomos[(r.n, r.theta)]
            return t.R(r.x) if r.op == "R" else t.RT(r.x)
        rec = rec_for(r.geometry)
        if isinstance(r, OperatorRequest):
            return (rec.cl_prop.D if r.op == "D" else rec.cl_prop.DT)(r.x, r.j)
        A = r.nangles
        if isinstance(r, GradientRequest):
            rec.data_size = r.data_size
            gr = {"prb": torch.zeros_like(r.prb)}
            for st in range(0, A, 16):
                B = min(16, A - st)
                sl = {"prb": r.prb, "proj": r.proj[st:st + B], "pos": r.pos[st:st + B]}
                g = {"prb": gr["prb"], "proj": torch.empty_like(sl["proj"]), "pos": torch.empty_like(sl["pos"])}
                rec._chunk(0, B)
                rec.gradients_cascade(sl, g, r.data[st:st + B])
            return gr
        rec.data_size = r.data_size
        tot = 0.0
        for st in range(0, A, 16):
            B = min(16, A - st)
            x = {"prb": r.x["prb"], "proj": r.x["proj"][st:st + B], "pos": r.x["pos"][st:st + B]}
            y = {"prb": r.y["prb"], "proj": r.y["proj"][st:st + B], "pos": r.y["pos"][st:st + B]}
            tot += rec.hessian_cascade(x, y, y, r.data[st:st + B])
        return tot

    plain_s = None
    if not a.no_plain:
        for r in reqs[:4]:
            plain(r)
        torch.cuda.synchronize()
        tp = time.perf_counter()
        for r in reqs:
            plain(r)
            torch.cuda.synchronize()
        plain_s = time.perf_counter() - tp
        say(f"plain {plain_s:.3f} s")
    counts = {k: sum(1 for sp in specs if sp[0] == k) for k in kinds}
    return {"metric": "Serving mix requests/s (served vs plain), p50/p99 latency", "value": round(nreq / cap_s, 3),
            "unit": "requests/s", "n_gpus": 1, "steps": nreq, "warmup": 2 * nreq,
            "ms_per_step": round(cap_s * 1e3 / nreq, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "c64", "data": "synthetic (seeded)",
            "config": {"workload": f"{nreq} requests; n in 256..2048, 1-4 distances, 16-128 angles; kinds operator (D/DT/R/RT) "
                                   f"/ gradient / bilinear = {counts}; value = all submitted at once; "
                                   f"latency at Poisson arrivals of {rate:g}/s"},
            "plain_value": round(nreq / plain_s, 3) if plain_s else None,
            "speedup_vs_plain": round(plain_s / cap_s, 3) if plain_s else None,
            "offered": {"rate": rate, "wall_s": round(wall_s, 3), "requests_per_s": round(nreq / wall_s, 3),
                        "latency_ms": {"p50": round(lat[len(lat) // 2], 3),
                                       "p99": round(lat[min(len(lat) - 1, int(0.99 * len(lat)))], 3)}},
            "capacity_s": round(cap_s, 3), "plain_s": round(plain_s, 3) if plain_s else None}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {a.gpus}",
              file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)
    if a.copy_calibration:
        src = torch.ones(1 << 25, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if a.op:
        print(json.dumps({"prop": op_prop, "tomo": op_tomo, "shift": op_shift, "bh": op_bh,
                          "serve-mix": op_serve_mix, "chunking": op_chunking}[a.op](a)), flush=True)
        return
    rec, vars_, data, host = make_inputs(a)
    grads = {"prb": torch.zeros_like(vars_["prb"]), "proj": torch.empty_like(vars_["proj"]),
             "pos": torch.empty_like(vars_["pos"])}
    path = "served" if a.path == "auto" else a.path
    fn_plain = lambda: plain_step(rec, vars_, grads, data)  # noqa: E731
    fused_value = None
    if path == "plain":
        wall, per = time_steps(fn_plain, a.steps, a.warmup)
        plain_value = value = a.angles * a.steps / wall
    else:
        plain_value = None
        if not a.no_plain:
            n_plain = max(2, a.steps // 2)
            wall_p, _ = time_steps(fn_plain, n_plain, 1)
            plain_value = a.angles * n_plain / wall_p
        fn = lambda: fused_step(rec, vars_, grads, data)  # noqa: E731
        fused_value = None
        if path == "fused":
            wall, per = time_steps(fn, a.steps, a.warmup)
        else:
            # the same fused launches called back to back, for comparison
            wall_f, _ = time_steps(fn, a.steps, a.warmup)
            fused_value = a.angles * a.steps / wall_f
            wall, per = served_run(a, rec, vars_, data, a.steps, a.warmup)
        value = a.angles * a.steps / wall
    ms = wall * 1e3 / a.steps
    roof = roofline_fresnel_cols(rec, a) if path == "plain" else roofline_fused(a, rec, vars_, grads, data)
    out = {
        "metric": "BH gradient evals/s at 2048^2 (served vs plain), % HBM roofline, p50/p99 latency",
        "value": round(value, 3), "unit": "evals/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64", "data": "synthetic (seeded phantom projections, Gaussian probe, noisy amplitudes)",
        "config": {"workload": f"BH gradient (gradients_cascade), detector {a.n}x{a.n}, {a.ndist} distances, "
                               f"object plane {a.nobj}x{a.nobj}, {a.angles} angles per request"
                               + (f", {a.steps} requests through the serving loop" if path == "served" else ""),
                   "path": path, "angles_per_step": a.angles, "n": a.n, "ndist": a.ndist, "nobj": a.nobj},
        "latency_ms": {"p50": round(per[len(per) // 2], 3), "p99": round(per[min(len(per) - 1, int(0.99 * len(per)))], 3)},
        "plain_value": round(plain_value, 3) if plain_value else None,
        "speedup_vs_plain": round(value / plain_value, 3) if plain_value else None,
        "roofline": roof,
    }
    if path == "served" and fused_value:
        out["fused_value"] = round(fused_value, 3)
        out["served_vs_fused"] = round(value / fused_value, 3)
    if isinstance(roof, dict):
        roof["measured_copy_gbs"] = measured_copy_gbs()  # beside the 8 TB/s spec peak
    if not a.no_cpu_baseline:
        out["cpu_baseline"], sample = cpu_baseline(a, host, vars_, data)
        out["parity_vs_fp64"] = parity_vs_fp64(rec, sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
