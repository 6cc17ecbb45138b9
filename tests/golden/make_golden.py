"""Golden vectors of the operator stack, written by the FP64 oracle (test
infrastructure: oracle/ is pinned by closed forms, adjoint identities and
finite differences in tests/test_oracle_*.py).

    python tests/golden/make_golden.py        # rewrites tests/golden/golden.npz

`cases()` returns {name: array}: seeded inputs and the oracle's outputs for
  * Fresnel D_0 and DT_1 of two 32x32 fields (propagation.py:70-112);
  * Radon R and RT of a 4-slice 32^3 slab at 12 angles, mask radius 0.9
    (tomo.py:56-116);
  * B-spline coeff, S and Sadj of 3 planes 40x48 sampled onto 30x32
    (shift.py:74-122);
  * the cascade objective, gradient (gradients_cascade) and bilinear form
    (hessian_cascade) of a 2-angle, 3-distance problem, detector 32^2, object
    40^2, complex and real objects (rec_mpi.py:250-353, 517-813).
Sizes are small so the file stays around half a megabyte.  tests/test_golden.py
checks that the oracle still reproduces the file (CPU) and that the HIP path
matches it (GPU).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import cascade as oc  # noqa: E402
from oracle import propagation as op  # noqa: E402
from oracle import shift as osh  # noqa: E402
from oracle import tomo as ot  # noqa: E402

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")

# physics of the reference's synthetic test (tests/test.py:14-19), 16x binned
ENERGY, FDD, PIXEL = 17.1, 1.217, 1.4760147601476e-6 * 16
Z1 = np.array([5.110, 5.464, 6.879, 9.817]) * 1e-3


def physics(ndist):
    """rec_mpi.py:55-60: wavelength, distances, voxel size, normalised magnifications."""
    z1 = Z1[:ndist]
    mags = FDD / z1
    norm = mags / mags[0]
    return 1.24e-09 / ENERGY, (z1 * (FDD - z1)) / FDD * norm ** 2, PIXEL / mags[0], norm


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def cases():
    g = {}
    # ---- Fresnel D / DT
    rng = np.random.default_rng(2001)
    wl, dist, vs, _ = physics(2)
    psi = crandn(rng, (2, 32, 32)).astype(np.complex64)
    prop = op.Propagation(32, 32, 2, 2, wl, vs, dist)
    g.update(prop_psi=psi, prop_wavelength=wl, prop_voxelsize=vs, prop_distance=dist,
             prop_D0=prop.D(psi.astype(np.complex128), 0), prop_DT1=prop.DT(psi.astype(np.complex128), 1))
    # ---- Radon R / RT
    rng = np.random.default_rng(2002)
    theta = np.arange(12) * np.pi / 12
    obj = crandn(rng, (4, 32, 32)).astype(np.complex64)
    sino = crandn(rng, (12, 4, 32)).astype(np.complex64)
    tomo = ot.Tomo(32, 4, theta, 0.9)
    g.update(tomo_theta=theta, tomo_obj=obj, tomo_sino=sino, tomo_R=tomo.R(obj.astype(np.complex128)),
             tomo_RT=tomo.RT(sino.astype(np.complex128)))
    # ---- B-spline coeff / S / Sadj
    rng = np.random.default_rng(2003)
    c = crandn(rng, (3, 40, 48)).astype(np.complex64)
    s = crandn(rng, (3, 30, 32)).astype(np.complex64)
    r = rng.uniform(-4, 4, (3, 2)).astype(np.float32)
    m = rng.uniform(0.9, 1.3, 3).astype(np.float32)
    sh = osh.Shift(32, 48, 30, 40, "complex64")
    g.update(shift_c=c, shift_s=s, shift_r=r, shift_m=m, shift_coeff=sh.coeff(c.astype(np.complex128)),
             shift_S=sh.S(c.astype(np.complex128), r, m), shift_Sadj=sh.Sadj(s.astype(np.complex128), r, m))
    # ---- cascade objective, gradient and bilinear form
    for real in (False, True):
        key = "cas_real" if real else "cas"
        rng = np.random.default_rng(2004 + real)
        n, nobj, nd, B = 32, 40, 3, 2
        wl, dist, vs, norm = physics(nd)
        dt = "float32" if real else "complex64"
        prop = op.Propagation(n, n, B, nd, wl, vs, dist)
        sh = osh.Shift(n, nobj, n, nobj, dt)
        rho = (1.0, 0.5, 0.2)
        cas = oc.Cascade(prop, sh, nd, n, n, nobj, nobj, B * nd * n * n, np.tile(1 / norm, (B, 1)), rho=rho,
                         obj_real=real)
        yy, xx = np.mgrid[0:n, 0:n] - n / 2
        prb = np.stack([np.exp(-(xx ** 2 + yy ** 2) / (2 * (0.4 * n) ** 2)) * np.exp(0.3j * np.cos(0.2 * xx + j))
                        for j in range(nd)]).astype(np.complex64)
        proj = (0.3 * (rng.standard_normal((B, nobj, nobj)) if real else crandn(rng, (B, nobj, nobj)))).astype(dt)
        pos = rng.uniform(-3, 3, (B, nd, 2)).astype(np.float32)
        d = (np.abs(cas.forward(prb, proj, pos)) * (1 + 0.05 * rng.standard_normal((B, nd, n, n)))).astype(np.float32)
        x = [(prb * (1 + 0.1 * crandn(rng, prb.shape))).astype(np.complex64),
             (proj + 0.05 * (rng.standard_normal(proj.shape) if real else crandn(rng, proj.shape))).astype(dt),
             (pos + rng.uniform(-0.5, 0.5, pos.shape)).astype(np.float32)]
        y = [(0.1 * crandn(rng, prb.shape)).astype(np.complex64),
             (0.1 * (rng.standard_normal(proj.shape) if real else crandn(rng, proj.shape))).astype(dt),
             (0.3 * rng.standard_normal(pos.shape)).astype(np.float32)]
        wide = [x[0].astype(np.complex128), x[1].astype(np.float64 if real else np.complex128),
                x[2].astype(np.float64)]
        ywide = [y[0].astype(np.complex128), y[1].astype(np.float64 if real else np.complex128),
                 y[2].astype(np.float64)]
        gprb, gproj, gpos = cas.gradients(*wide, d)
        g.update({f"{key}_prb": x[0], f"{key}_proj": x[1], f"{key}_pos": x[2], f"{key}_data": d,
                  f"{key}_yprb": y[0], f"{key}_yproj": y[1], f"{key}_ypos": y[2],
                  f"{key}_objective": cas.objective(*wide, d),
                  f"{key}_gprb": gprb, f"{key}_gproj": gproj, f"{key}_gpos": gpos,
                  f"{key}_hess_yy": cas.hessian(wide, ywide, ywide, d)})
    return g


def stored(g):
    """What the file keeps: complex128 outputs rounded to complex64 (the GPU
    checks are at 1e-5 or looser); scalars and real arrays as computed."""
    return {k: (v.astype(np.complex64) if np.iscomplexobj(v) and v.dtype == np.complex128 else v)
            for k, v in g.items()}


if __name__ == "__main__":
    np.savez_compressed(PATH, **stored(cases()))
    print(f"wrote {PATH} ({os.path.getsize(PATH)} bytes)")
