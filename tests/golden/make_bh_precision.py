"""Ten FP64 oracle iterations of the BH loop on the precision-study problem
(bh_precision_case.py), written to tests/golden/bh_precision.npz:
per-iteration (alpha, beta, objective), the final positions, the final
object's central slice and norm, and every fourth row and column of the final
probe.

    python tests/golden/make_bh_precision.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

import bh_precision_case as case  # noqa: E402


def main():
    from oracle import propagation as op
    from oracle import shift as osh
    from oracle import tomo as ot
    op.WORKERS = osh.WORKERS = ot.WORKERS = os.cpu_count()
    t0 = time.time()
    _, orc, _, _, _, start64 = case.problem()
    print(f"problem built in {time.time() - t0:.0f} s", flush=True)
    trace = []
    vo, hist = orc.BH(start64, case.NITER, trace=trace)
    print(f"{case.NITER} iterations in {time.time() - t0:.0f} s", flush=True)
    for i, (a, b, e) in enumerate(hist):
        print(f"iter {i}: alpha {a:.6e} beta {b:.6e} objective {e:.9e}", flush=True)
    out = os.path.join(HERE, "bh_precision.npz")
    np.savez_compressed(out, hist=np.asarray(hist, dtype=np.float64), pos=vo["pos"],
                        obj_mid=vo["obj"][case.NOBJ // 2].astype(np.complex128),
                        obj_norm=np.float64(np.linalg.norm(vo["obj"])),
                        prb_sub=vo["prb"][:, ::4, ::4].astype(np.complex64))
    print("wrote", out, os.path.getsize(out), "bytes", flush=True)


if __name__ == "__main__":
    main()
