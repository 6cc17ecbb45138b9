"""Precision study of the BH loop at a reduced shape of one rank's share of
the largest configuration (detector 256^2, 4 distances, object 128^3 of
random ellipsoids, 64 angles, probe fit and Laplacian regulariser on): ten
FP32 iterations on the GPU (fused gradient and bilinear forms, FP64 tables,
FP64 reductions) against ten FP64 iterations of the oracle
(tests/golden/bh_precision.npz, written by tests/golden/make_bh_precision.py
from the same seeded problem, tests/golden/bh_precision_case.py).

Bars: objective after every iteration 1e-4 relative; step sizes alpha 1e-3
and beta 1e-2 relative or 1e-3 absolute (beta is a ratio whose numerator
nearly cancels); final variables 1e-3 relative L2 (step-size differences
accumulate along the iterations)."""
import os
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.linalg.norm(a.astype(np.complex128) - b) / np.linalg.norm(b))


def test_bh_ten_iterations_reduced_share_vs_fp64(gpu):
    if GOLDEN not in sys.path:
        sys.path.insert(0, GOLDEN)
    import bh_precision_case as case
    from holotomocupy_amd.rec import Rec
    from oracle import propagation as op
    from oracle import tomo as ot
    op.WORKERS = ot.WORKERS = 16
    print("[case] BH precision study 256^2 / 128^3 / 64 angles / 4 distances, 10 iterations", flush=True)
    ref = np.load(os.path.join(GOLDEN, "bh_precision.npz"))
    args, _, data, refamp, start, _ = case.problem()
    rec = Rec(args)
    rec.data.copy_(torch.from_numpy(data))
    rec.ref.copy_(torch.from_numpy(refamp))
    for k in ("obj", "prb", "pos"):
        rec.vars[k].copy_(torch.from_numpy(start[k]))
    rec.BH()
    torch.cuda.synchronize()
    hist = ref["hist"]
    errs = [e for i, e, _ in rec.table if i >= 0]
    obj_rel = [abs(a - h) / abs(h) for a, h in zip(errs, hist[:, 2])]
    print("[precision] objective rel. error per iteration", ["%.2e" % r for r in obj_rel], flush=True)
    print("[precision] alpha rel. error", ["%.1e" % (abs(float(a) - h) / abs(h)) for a, h in zip(rec.alphas, hist[:, 0])],
          flush=True)
    assert len(errs) == case.NITER and max(obj_rel) < 1e-4
    for k in range(case.NITER):
        assert float(rec.alphas[k]) == pytest.approx(hist[k, 0], rel=1e-3)
        assert float(rec.betas[k]) == pytest.approx(hist[k, 1], rel=1e-2, abs=1e-3)
    e_pos = rel(rec.vars["pos"], ref["pos"])
    e_obj = rel(rec.vars["obj"][case.NOBJ // 2], ref["obj_mid"])
    e_prb = rel(rec.vars["prb"][:, ::4, ::4], ref["prb_sub"])
    print(f"[precision] final variables rel. L2: obj (central slice) {e_obj:.2e}, prb {e_prb:.2e}, pos {e_pos:.2e}",
          flush=True)
    assert e_obj < 1e-3 and e_prb < 1e-3 and e_pos < 1e-3
