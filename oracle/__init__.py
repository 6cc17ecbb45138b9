"""CPU oracle for the BH operator stack (TEST INFRASTRUCTURE ONLY).

This package restates, in NumPy/SciPy, the algorithm of the reference
operators that sit under the Bilinear-Hessian solver of holotomocupy:

  * ``oracle.propagation`` -- Fresnel propagator D / DT
    (src/holotomocupy/propagation.py:14-112, cuda_kernels.py:78-127)
  * ``oracle.shift``       -- cubic B-spline shift family S / Sadj / dS / d2S / dSadj
    and the periodic prefilter ``coeff`` (src/holotomocupy/shift.py:18-221,
    cuda_kernels.py:131-843)
  * ``oracle.tomo``        -- USFFT Radon transform R / RT
    (src/holotomocupy/tomo.py:12-116, cuda_kernels.py:3-76)
  * ``oracle.cascade``     -- the F0..F3 cascade, its gradient and its bilinear
    (Hessian) form (src/holotomocupy/rec_mpi.py:250-353, 517-813)

Every function defaults to float64 / complex128 and can run in float32 /
complex64 for the CPU baseline timing.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this package, and only as the checker / CPU baseline.  The
product package ``holotomocupy_amd`` never imports it.

Parity pins: every function here is checked in ``tests/test_oracle_*.py``
against closed forms, library routines (scipy.ndimage, DCT, brute-force
NUDFT) or invariants chosen so that a dropped term, a wrong sign or a
transposed index fails; the per-function pins are listed in each module
docstring.  The reference itself cannot be imported here (cupy and an
NVIDIA GPU are absent), so no function is pinned against reference outputs.
"""
