"""Serving loop: scheduling core on the CPU (fake executor), result identity
alone vs in the mix and graph replay on the GPU."""
import numpy as np
import pytest

from holotomocupy_amd.serve import BHServer, Geometry, GradientRequest, Scheduler


class _T:
    """Stand-in for a device tensor: only .shape and .dtype are used by the
    scheduler."""
    dtype = "complex64"

    def __init__(self, *shape):
        self.shape = shape


class FakeExec:
    def __init__(self, per_angle=100):
        self.per_angle = per_angle
        self.runs = []
        self.pending_done = set()

    def footprint(self, req):
        return req.nangles * self.per_angle

    def run(self, geo, batch):
        h = len(self.runs)
        self.runs.append((geo, [r.nangles for r in batch]))
        return h

    def done(self, h):
        return h in self.pending_done

    def finish(self, h, batch):
        for r in batch:
            r.future.set_result(h)


def req(geo, a):
    return GradientRequest(geo, _T(1), _T(a, 4, 4), _T(a, 1, 2), _T(a, 1, 4, 4), data_size=1.0)


G1 = Geometry(8, 8, 8, 8, 1)
G2 = Geometry(16, 16, 16, 16, 1)


def test_admission_by_budget_and_release():
    ex = FakeExec()
    s = Scheduler(ex, hbm_budget=3000, max_batch_angles=64, max_inflight=1)
    rs = [req(G1, 16) for _ in range(3)]  # 1600 bytes each
    for r in rs:
        s.submit(r)
    s.step()
    assert s.in_use == 1600 and len(s.pending) == 2 and ex.runs == [(G1, [16])]
    ex.pending_done.add(0)
    s.step()
    assert rs[0].future.result() == 0
    assert len(ex.runs) == 2 and len(s.pending) == 1
    ex.pending_done.update({1, 2})
    for _ in range(3):
        s.step()
    assert all(r.future.done() for r in rs) and s.idle() and s.in_use == 0


def test_rejects_oversize_requests():
    ex = FakeExec()
    s = Scheduler(ex, hbm_budget=1000, max_batch_angles=8)
    with pytest.raises(MemoryError):
        s.submit(req(G1, 16)).result()
    with pytest.raises(ValueError):
        Scheduler(ex, 10 ** 9, max_batch_angles=8).submit(req(G1, 16)).result()


def test_batches_per_geometry_fifo_and_fair():
    ex = FakeExec(per_angle=1)
    s = Scheduler(ex, hbm_budget=10 ** 9, max_batch_angles=64, max_inflight=4)
    order = [(G1, 32), (G2, 16), (G1, 32), (G1, 16), (G2, 64), (G1, 8)]
    for g, a in order:
        s.submit(req(g, a))
    s.step()
    # first bucket: G1 gets 32 + 32 (16 would exceed 64); then G2: 16 (64 would exceed); buckets alternate
    assert ex.runs[0] == (G1, [32, 32])
    assert ex.runs[1] == (G2, [16])
    assert ex.runs[2] == (G1, [16, 8])
    assert ex.runs[3] == (G2, [64])


def test_server_thread_with_fake_executor():
    ex = FakeExec()
    srv = BHServer(hbm_budget=10 ** 9, executor=ex)
    futs = [srv.submit(req(G1, 8)) for _ in range(5)]
    import time
    t0 = time.time()
    while len(ex.runs) < 1 and time.time() - t0 < 5:
        time.sleep(0.001)
    ex.pending_done.update(range(100))
    assert all(f.result(timeout=5) is not None for f in futs)
    srv.close()


# ------------------------------------------------------------------ GPU
def _gpu_request(rng, geo, a, data_size, rho, scale):
    import torch
    from holotomocupy_amd import phantom
    prb = torch.from_numpy(phantom.probe(rng, geo.ndist, geo.nz, geo.n) * scale).cuda()
    proj = torch.from_numpy(phantom.projections(rng, a, geo.nzobj, geo.nobj)).cuda()
    pos = torch.from_numpy(phantom.positions(rng, a, geo.ndist, 2.0)).cuda()
    data = torch.from_numpy(rng.uniform(0.5, 1.5, (a, geo.ndist, geo.nz, geo.n)).astype(np.float32)).cuda()
    return GradientRequest(geo, prb, proj, pos, data, data_size=data_size, rho=rho)


@pytest.mark.gpu
@pytest.mark.parametrize("graphs", [False, True])
def test_result_same_alone_and_in_mix(gpu, graphs):
    import torch
    print(f"[case] serve alone vs mix graphs={graphs}", flush=True)
    rng = np.random.default_rng(77)
    ga = Geometry(64, 64, 48, 48, 3, detector_pixelsize=1.476e-6 * 64)
    gb = Geometry(32, 32, 40, 40, 2, detector_pixelsize=1.476e-6 * 128)
    specs = [(ga, 4, 1e6, (1, 0.5, 0.2), 1.0), (gb, 3, 2e5, (1, 1, 1), 0.7), (ga, 5, 3e6, (2, 0.1, 0.3), 1.3),
             (ga, 2, 1e6, (1, 1, 1), 0.9)]
    reqs = [_gpu_request(np.random.default_rng(100 + i), *sp) for i, sp in enumerate(specs)]
    srv = BHServer(hbm_budget=8 << 30, use_graphs=graphs)
    alone = []
    for r in reqs:
        alone.append([t.clone() for t in srv.submit(GradientRequest(r.geometry, r.prb, r.proj, r.pos, r.data,
                                                                     r.data_size, r.rho)).result(timeout=60)])
    mixed = [srv.submit(GradientRequest(r.geometry, r.prb, r.proj, r.pos, r.data, r.data_size, r.rho))
             for r in reqs]
    for _ in range(3):  # replays with eager work in between
        x = torch.randn(1000, device="cuda")
        torch.cuda.synchronize()
        again = [srv.submit(GradientRequest(r.geometry, r.prb, r.proj, r.pos, r.data, r.data_size, r.rho))
                 for r in reqs]
        for a, f in zip(alone, again):
            for u, v in zip(a, f.result(timeout=60)):
                assert torch.equal(u, v)
        del x
    for a, f in zip(alone, mixed):
        for u, v in zip(a, f.result(timeout=60)):
            assert torch.equal(u, v)
    srv.close()


def test_mixed_kinds_are_bucketed_by_key():
    from holotomocupy_amd.serve import OperatorRequest
    ex = FakeExec(per_angle=1)
    s = Scheduler(ex, hbm_budget=10 ** 9, max_batch_angles=64, max_inflight=8)
    s.submit(req(G1, 16))
    s.submit(OperatorRequest(G1, "D", 0, _T(8, 8, 8)))
    s.submit(OperatorRequest(G1, "D", 1, _T(4, 8, 8)))
    s.submit(OperatorRequest(G1, "D", 0, _T(8, 8, 8)))
    s.submit(OperatorRequest(G1, "DT", 0, _T(2, 8, 8)))
    s.step()
    keys = [k for k, _ in ex.runs]
    assert keys == [G1, (G1, "prop", "D", 0), (G1, "prop", "D", 1), (G1, "prop", "DT", 0)]
    assert ex.runs[1][1] == [8, 8]  # the two D_0 requests share a launch


@pytest.mark.gpu
def test_operator_and_bilinear_requests(gpu):
    """Propagation requests packed into one launch equal the direct D / DT
    calls bit for bit; bilinear-form requests equal Rec.hessian_fused; mixed
    with a gradient request."""
    import torch
    from holotomocupy_amd.serve import BilinearRequest, OperatorRequest
    print("[case] serve operator and bilinear requests", flush=True)
    rng = np.random.default_rng(78)
    ga = Geometry(64, 64, 48, 48, 3, detector_pixelsize=1.476e-6 * 64)
    srv = BHServer(hbm_budget=8 << 30)
    g = torch.Generator(device="cuda").manual_seed(78)
    xs = [torch.randn((b, 64, 64), dtype=torch.complex64, device="cuda", generator=g) for b in (3, 5, 2)]
    ops = [OperatorRequest(ga, "D", 1, xs[0]), OperatorRequest(ga, "DT", 2, xs[1]), OperatorRequest(ga, "D", 1, xs[2])]
    greq = _gpu_request(rng, ga, 4, 1e6, (1, 0.5, 0.2), 1.0)
    r = _gpu_request(rng, ga, 3, 1e6, (1, 1, 1), 1.0)
    xdir = {"prb": r.prb, "proj": r.proj, "pos": r.pos}
    ydir = {"prb": 0.1 * r.prb, "proj": 0.05 * torch.randn_like(r.proj), "pos": 0.3 * torch.randn_like(r.pos)}
    hreqs = [BilinearRequest(ga, xdir, ydir, None, r.data, 1e6), BilinearRequest(ga, xdir, ydir, xdir, r.data, 1e6)]
    futs = [srv.submit(q) for q in ops] + [srv.submit(greq)] + [srv.submit(q) for q in hreqs]
    res = [f.result(timeout=60) for f in futs]
    b = srv.sched.ex.buckets[ga]
    prop = b["rec"].cl_prop
    assert torch.equal(res[0], prop.D(xs[0], 1))
    assert torch.equal(res[1], prop.DT(xs[1], 2))
    assert torch.equal(res[2], prop.D(xs[2], 1))
    mag = b["rec"].eff_demagnifications[0].expand(3, -1).contiguous()
    for q, got in zip(hreqs, res[4:]):
        ref = torch.zeros(1, dtype=torch.float64, device="cuda")
        b["bh"].hessian(q.x, q.y, q.y if q.z is None else q.z, q.data, mag, ref, accumulate=True, scale=2.0 / 1e6)
        assert float(got) == float(ref)
    srv.close()


class _FailingExec(FakeExec):
    """Raises in run() for batches of one geometry (e.g. device memory
    exhausted)."""
    def run(self, geo, batch):
        if geo == G2:
            raise MemoryError("device memory exhausted")
        return super().run(geo, batch)


def test_failed_batch_fails_its_requests_and_the_loop_goes_on():
    ex = _FailingExec(per_angle=1)
    srv = BHServer(hbm_budget=10 ** 9, executor=ex)
    bad = srv.submit(req(G2, 8))
    good = srv.submit(req(G1, 8))
    with pytest.raises(MemoryError):
        bad.result(timeout=5)
    ex.pending_done.update(range(100))
    assert good.result(timeout=5) is not None
    srv.close()
    assert srv.sched.in_use == 0  # the failed request's bytes were released too


@pytest.mark.gpu
def test_arena_growth_and_graph_capture_keep_results(gpu):
    """Geometries of growing size share each stream's arena (it grows, its
    graphs are dropped); shapes seen before run from captured graphs.  Every
    result equals the eager server's bit for bit."""
    import torch
    from holotomocupy_amd.serve import GpuExecutor
    print("[case] serve arena growth and graph capture", flush=True)
    gb = Geometry(32, 32, 40, 40, 2, detector_pixelsize=1.476e-6 * 128)
    ga = Geometry(64, 64, 48, 48, 3, detector_pixelsize=1.476e-6 * 64)
    specs = [(gb, 3, 2e5, (1, 1, 1), 0.7), (ga, 4, 1e6, (1, 0.5, 0.2), 1.0), (gb, 3, 2e5, (1, 1, 1), 0.7),
             (ga, 4, 1e6, (1, 0.5, 0.2), 1.0), (ga, 4, 1e6, (1, 0.5, 0.2), 1.0), (gb, 3, 2e5, (1, 1, 1), 0.7)]
    reqs = [_gpu_request(np.random.default_rng(200 + i), *sp) for i, sp in enumerate(specs)]

    def serve(graphs):
        ex = GpuExecutor(use_graphs=graphs, nstreams=1)
        srv = BHServer(hbm_budget=1 << 30, executor=ex)
        out = []
        for r in reqs:  # one at a time: every batch is one request
            q = GradientRequest(r.geometry, r.prb, r.proj, r.pos, r.data, r.data_size, r.rho)
            out.append(srv.submit(q).result(timeout=60))
        srv.close()
        return out, ex

    ref, _ = serve(False)
    got, ex = serve(True)
    assert ex.arena_growths == 2 and len(ex.slots[0]["graphs"]) == 2
    for a, b in zip(ref, got):
        for x, y in zip(a, b):
            assert torch.equal(x, y)


def test_tomo_requests_bucket_by_grid_angles_and_operator():
    from holotomocupy_amd.serve import TomoRequest
    ex = FakeExec(per_angle=1)
    s = Scheduler(ex, hbm_budget=10 ** 9, max_batch_angles=64, max_inflight=8)
    th = (0.0, 0.5, 1.0)
    s.submit(TomoRequest(8, th, 0.0, "R", _T(4, 8, 8)))
    s.submit(TomoRequest(8, th, 0.0, "RT", _T(3, 2, 8)))
    s.submit(TomoRequest(8, th, 0.0, "R", _T(2, 8, 8)))
    s.submit(TomoRequest(8, (0.0, 0.5), 0.0, "R", _T(1, 8, 8)))
    s.step()
    assert [b for _, b in ex.runs] == [[4, 2], [2], [1]]  # slices per batch


@pytest.mark.gpu
def test_tomo_requests_equal_direct_calls(gpu):
    """R / RT requests stacked along z in one launch equal Tomo.R / RT of
    each slab alone bit for bit."""
    import torch
    from holotomocupy_amd.serve import TomoRequest
    from holotomocupy_amd.tomo import Tomo
    print("[case] serve tomo requests", flush=True)
    n, nth = 64, 24
    theta = tuple(float(t) for t in np.arange(nth) * np.pi / nth)
    g = torch.Generator(device="cuda").manual_seed(79)
    objs = [torch.randn((k, n, n), dtype=torch.complex64, device="cuda", generator=g) for k in (3, 5)]
    sinos = [torch.randn((nth, k, n), dtype=torch.complex64, device="cuda", generator=g) for k in (2, 4)]
    srv = BHServer(hbm_budget=8 << 30)
    reqs = [TomoRequest(n, theta, 0.9, "R", o) for o in objs] + [TomoRequest(n, theta, 0.9, "RT", d) for d in sinos]
    futs = [srv.submit(r) for r in reqs]
    res = [f.result(timeout=60) for f in futs]
    srv.close()
    t = Tomo(n, 16, np.asarray(theta), 0.9)
    for o, got in zip(objs, res[:2]):
        assert torch.equal(got, t.R(o))
    for d, got in zip(sinos, res[2:]):
        assert torch.equal(got, t.RT(d))


def test_submit_many_forms_the_same_batches():
    ex = FakeExec(per_angle=1)
    srv = BHServer(hbm_budget=10 ** 9, executor=ex)
    for _ in range(3):
        futs = srv.submit_many([req(G1, 16) for _ in range(10)])
        ex.pending_done.update(range(1000))
        for f in futs:
            f.result(timeout=5)
    srv.close()
    # every burst: the scheduler sees all 10 at once and packs them the same way
    sizes = [b for _, b in ex.runs]
    per_burst = len(sizes) // 3
    assert sizes[:per_burst] == sizes[per_burst:2 * per_burst] == sizes[2 * per_burst:]
