"""Committed golden vectors (tests/golden/golden.npz, written by
tests/golden/make_golden.py from the FP64 oracle).

CPU: the oracle still reproduces the file (inputs bit for bit, outputs to the
complex64 rounding of the file, relative L2 <= 1e-6; scalars to 1e-10).
GPU: the HIP path matches the file: relative L2 <= 1e-5 for fields and the
object / probe gradients, 1e-4 for the position gradient and the scalars
(sums of many signed FP32 terms)."""
import os
import sys
from types import SimpleNamespace

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402

GOLDEN = np.load(os.path.join(HERE, "golden", "golden.npz"))


def rel(a, b):
    a = a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a.astype(np.complex128) - b) / np.linalg.norm(b))


OUTPUTS = ("D0", "DT1", "R", "RT", "coeff", "S", "Sadj", "gprb", "gproj", "gpos", "objective", "yy", "data")


def test_oracle_reproduces_golden():
    """Seeded inputs bit for bit; oracle outputs (FFT rounding may differ with
    the host's SIMD paths) to the file's complex64 rounding."""
    now = make_golden.stored(make_golden.cases())
    assert sorted(now) == sorted(GOLDEN.files)
    for k in GOLDEN.files:
        want, got = GOLDEN[k], now[k]
        if k.split("_")[-1] not in OUTPUTS:
            np.testing.assert_array_equal(got, want, err_msg=k)
        elif want.ndim == 0:
            assert float(got) == pytest.approx(float(want), rel=1e-10), k
        else:
            assert rel(got, want) <= 1e-6, k


# ------------------------------------------------------------------ GPU
def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.gpu
def test_operators_match_golden(gpu):
    from holotomocupy_amd import Propagation, Shift, Tomo
    print("[case] golden D/DT, R/RT, coeff/S/Sadj", flush=True)
    g = GOLDEN
    prop = Propagation(32, 32, 2, 2, float(g["prop_wavelength"]), float(g["prop_voxelsize"]), g["prop_distance"])
    assert rel(prop.D(dev(g["prop_psi"]), 0), g["prop_D0"]) < 1e-5
    assert rel(prop.DT(dev(g["prop_psi"]), 1), g["prop_DT1"]) < 1e-5
    tomo = Tomo(32, 4, g["tomo_theta"], 0.9)
    assert rel(tomo.R(dev(g["tomo_obj"])), g["tomo_R"]) < 1e-5
    assert rel(tomo.RT(dev(g["tomo_sino"])), g["tomo_RT"]) < 1e-5
    sh = Shift(32, 48, 30, 40, "complex64")
    r, m = dev(g["shift_r"]), dev(g["shift_m"])
    assert rel(sh.coeff(dev(g["shift_c"])), g["shift_coeff"]) < 1e-5
    assert rel(sh.S(dev(g["shift_c"]), r, m), g["shift_S"]) < 1e-5
    assert rel(sh.Sadj(dev(g["shift_s"]), r, m), g["shift_Sadj"]) < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("real", [False, True])
def test_cascade_matches_golden(gpu, real):
    import torch
    from holotomocupy_amd.rec import Rec
    print(f"[case] golden cascade real={real}", flush=True)
    key = "cas_real" if real else "cas"
    g = {k[len(key) + 1:]: GOLDEN[k] for k in GOLDEN.files
         if k.startswith(key + "_") and (real or not k.startswith("cas_real"))}
    B, nd, n, _ = g["data"].shape
    nobj = g["proj"].shape[-1]
    args = SimpleNamespace(energy=make_golden.ENERGY, detector_pixelsize=make_golden.PIXEL,
                           focustodetectordistance=make_golden.FDD, z1=make_golden.Z1[:nd], ndist=nd, ntheta=B,
                           nz=n, n=n, nzobj=nobj, nobj=nobj, obj_dtype="float32" if real else "complex64",
                           rho=[1.0, 0.5, 0.2])
    rec = Rec(args)
    x = {"prb": dev(g["prb"]), "proj": dev(g["proj"]), "pos": dev(g["pos"])}
    y = {"prb": dev(g["yprb"]), "proj": dev(g["yproj"]), "pos": dev(g["ypos"])}
    d = dev(g["data"])
    assert rec.objective(x, d) == pytest.approx(float(g["objective"]), rel=1e-4)
    for path in ("fused", "cascade"):
        gr = {"prb": torch.zeros_like(x["prb"]), "proj": torch.empty_like(x["proj"]),
              "pos": torch.empty_like(x["pos"])}
        getattr(rec, f"gradients_{path}")(x, gr, d)
        assert rel(gr["prb"], g["gprb"]) < 1e-5, path
        assert rel(gr["proj"], g["gproj"]) < 1e-5, path
        assert rel(gr["pos"], g["gpos"]) < 1e-4, path
    assert float(rec.hessian_fused(x, y, y, d)) == pytest.approx(float(g["hess_yy"]), rel=1e-4)
    assert rec.hessian_cascade(x, y, y, d) == pytest.approx(float(g["hess_yy"]), rel=1e-4)
