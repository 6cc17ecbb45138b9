"""Oracle: the BH reconstruction loop of rec_mpi.Rec on one rank (TEST INFRASTRUCTURE ONLY).

Restates, in NumPy float64 / complex128, for a single process (the MPI
redistribution is the identity, all-reduces are the identity and the ghost
slices of the Laplacian are the zero slices of the global boundary):

  * the 3-D Laplacian of the regulariser, rec_mpi.py:430-436 (`_lap`: 7-point
    stencil, periodic in y and x, with the two z neighbours taken from a volume
    zero-padded by two slices on each side, rec_mpi.py:402-416);
  * gradient_laplacian (rec_mpi.py:418-443): g += 2 lam/N lap(lap(u_pad));
  * hessian_laplacian (rec_mpi.py:445-469): 2 lam/N Re<d1, lap(lap(e_pad))>;
  * _lap_energy_local (rec_mpi.py:471-494): lam/N ||lap(u)||^2 over the
    volume's own slices;
  * gradient_prbfit / hessian_prbfit (rec_mpi.py:364-398) and the probe-fit
    term of Rec.min (rec_mpi.py:830-835);
  * gradients (rec_mpi.py:305-325), hessian (rec_mpi.py:235-248), fwd_tomo
    (rec_mpi.py:498-507), gF4 (rec_mpi.py:355-362), min (rec_mpi.py:815-836),
    gen_sqrt_data / gen_sqrt_ref (rec_mpi.py:882-902);
  * BH (rec_mpi.py:138-232): the Hessian-conjugate direction update
    beta = H(g, eta)/H(eta, eta), the step alpha = -<grad f, eta>/H(eta, eta),
    and the update of obj, prb, pos and proj.

The cascade itself (data term, its gradient and bilinear form) is
oracle.cascade.Cascade; R / RT are oracle.tomo.Tomo.

Pins (tests/test_oracle_bh.py):
  * lap equals scipy.ndimage.laplace with constant (zero) z and wrapped y, x
    boundaries; the biharmonic of the gradient equals
    laplace(laplace(u zero-padded by two slices))[2:-2] and is the gradient of
    lam/N ||lap(u padded)||^2 over the padded volume (central differences);
  * the step alpha of the first iteration is the Newton step along eta of the
    objective: -phi'(0)/phi''(0) with phi(t) = min(x + t eta) from central
    differences of the objective (lam_laplacian = 0 there, since the reference's
    energy omits the ghost-slice Laplacians that its gradient includes);
  * the updated direction is Hessian-conjugate to the previous one in the
    bilinear form (H(eta_new, eta_old) = 0, each H from the pinned cascade);
  * the objective decreases over BH iterations on a synthetic problem.
"""
import numpy as np

from .cascade import Cascade, redot
from .propagation import Propagation
from .shift import Shift
from .tomo import Tomo


def lap(a, b, c):
    """rec_mpi.py:430-436: a, c are the z-1 and z+1 slices of b's slices."""
    return (a + c + np.roll(b, -1, axis=1) + np.roll(b, 1, axis=1)
            + np.roll(b, -1, axis=2) + np.roll(b, 1, axis=2) - 6 * b)


def pad2(u):
    """Volume with two zero ghost slices on each side (u_pad of rec_mpi.py:120-125,
    ghosts as set at the global boundary by rec_mpi.py:415-416)."""
    p = np.zeros((u.shape[0] + 4,) + u.shape[1:], dtype=u.dtype)
    p[2:-2] = u
    return p


def biharmonic(u):
    """lap(lap(u_pad)) on the volume's slices (rec_mpi.py:432-436)."""
    p = pad2(u)
    lzm1 = lap(p[:-4], p[1:-3], p[2:-2])
    lz = lap(p[1:-3], p[2:-2], p[3:-1])
    lzp1 = lap(p[2:-2], p[3:-1], p[4:])
    return lap(lzm1, lz, lzp1)


def laplacian_energy(u, lam, obj_size):
    """_lap_energy_local (rec_mpi.py:471-494)."""
    if lam == 0:
        return 0.0
    p = pad2(u)
    L = lap(p[1:-3], p[2:-2], p[3:-1])
    return lam / obj_size * redot(L, L)


class BHSolver:
    """Rec.BH and what it calls, for one rank holding every angle and slice."""

    def __init__(self, n, nz, nobj, nzobj, ndist, wavelength, voxelsize, distance, theta, mag,
                 rho=(1.0, 1.0, 1.0), mask_r=0, lam_prbfit=0.0, lam_laplacian=0.0, obj_real=False):
        self.n, self.nz, self.nobj, self.nzobj, self.ndist = n, nz, nobj, nzobj, ndist
        self.theta = np.asarray(theta, dtype=np.float64)
        self.ntheta = len(self.theta)
        self.prop = Propagation(n, nz, self.ntheta, ndist, wavelength, voxelsize, distance)
        self.shift = Shift(n, nobj, nz, nzobj, "float32" if obj_real else "complex64")
        self.tomo = Tomo(nobj, nzobj, self.theta, mask_r)
        self.data_size = self.ntheta * ndist * nz * n
        self.prb_size = ndist * nz * n
        self.obj_size = nzobj * nobj ** 2
        self.cas = Cascade(self.prop, self.shift, ndist, nz, n, nzobj, nobj, self.data_size, mag, rho, obj_real)
        self.rho_sq = self.cas.rho_sq
        self.norm_const = np.float32(np.sqrt(nobj / self.ntheta))
        self.lam_prbfit, self.lam_laplacian = lam_prbfit, lam_laplacian
        self.ref = None
        self.data = None
        self.obj_real = obj_real

    # ---- regulariser and probe fit
    def gradient_laplacian(self, grad_obj, u):
        if self.lam_laplacian == 0:
            return grad_obj
        return grad_obj + 2.0 * self.lam_laplacian / self.obj_size * biharmonic(u)

    def hessian_laplacian(self, dobj1, e):
        if self.lam_laplacian == 0:
            return 0.0
        return 2.0 * self.lam_laplacian / self.obj_size * redot(dobj1, biharmonic(e))

    def gradient_prbfit(self, prb):
        """rec_mpi.py:364-378, returned (already times rho_prb^2)."""
        g = np.zeros_like(prb)
        if self.lam_prbfit == 0:
            return g
        for j in range(self.ndist):
            tmp = self.prop.D(prb[j:j + 1], j)
            td = self.ref[j:j + 1] * (tmp / np.abs(tmp))
            g[j:j + 1] = self.lam_prbfit / self.prb_size * self.prop.DT(2 * (tmp - td), j) * self.rho_sq["prb"]
        return g

    def hessian_prbfit(self, prb, dprb1, dprb2):
        """rec_mpi.py:380-398."""
        if self.lam_prbfit == 0:
            return 0.0
        out = 0.0
        for j in range(self.ndist):
            Dp = self.prop.D(prb[j:j + 1], j)
            D1 = self.prop.D(dprb1[j:j + 1], j)
            D2 = self.prop.D(dprb2[j:j + 1], j)
            l0 = Dp / np.abs(Dp)
            d0 = self.ref[j:j + 1] / np.abs(Dp)
            rp = lambda a, b: np.real(a) * np.real(b) + np.imag(a) * np.imag(b)  # noqa: E731
            v1 = np.sum((1 - d0) * rp(D1, D2))
            v2 = np.sum(d0 * rp(l0, D1) * rp(l0, D2))
            out += 2 * (v1 + v2)
        return self.lam_prbfit * out / self.prb_size

    # ---- tomography
    def fwd_tomo(self, obj):
        return self.tomo.R(obj)

    def gF4(self, gradproj):
        return self.tomo.RT(gradproj)

    # ---- drivers (rec_mpi.py:235-325, 815-836)
    def gradients(self, vars):
        gprb, gproj, gpos = self.cas.gradients(vars["prb"], vars["proj"], vars["pos"], self.data)
        gobj = self.gF4(gproj)
        gobj = self.gradient_laplacian(gobj, vars["obj"])
        gprb = gprb + self.gradient_prbfit(vars["prb"])
        return {"obj": gobj, "prb": gprb, "pos": gpos, "proj": gproj}

    def hessian(self, vars, grads, etas):
        x = [vars["prb"], vars["proj"], vars["pos"]]
        y = [grads["prb"], grads["proj"], grads["pos"]]
        z = y if grads is etas else [etas["prb"], etas["proj"], etas["pos"]]
        w = self.cas.hessian(x, y, z, self.data)
        w += self.hessian_prbfit(vars["prb"], grads["prb"], etas["prb"])
        w += self.hessian_laplacian(grads["obj"], etas["obj"])
        return w

    def min(self, vars):
        out = self.cas.objective(vars["prb"], vars["proj"], vars["pos"], self.data)
        for j in range(self.ndist):
            Dprb = self.prop.D(vars["prb"][j:j + 1], j)[0]
            out += self.lam_prbfit / self.prb_size * np.linalg.norm(np.abs(Dprb) - self.ref[j]) ** 2
        return out + laplacian_energy(vars["obj"], self.lam_laplacian, self.obj_size)

    def gen_sqrt_data(self, vars):
        """rec_mpi.py:882-896 (obj given unnormalised, as there)."""
        proj = self.fwd_tomo(vars["obj"] / self.norm_const)
        return np.abs(self.cas.forward(vars["prb"], proj, vars["pos"]))

    def gen_sqrt_ref(self, prb):
        return np.stack([np.abs(self.prop.D(prb[j:j + 1], j)[0]) for j in range(self.ndist)])

    def BH(self, vars, niter, start_iter=0, mask=None, trace=None):
        """rec_mpi.py:138-232 on one rank.  vars holds obj (unnormalised),
        prb, pos; returns (vars, history) with history[i] = (alpha, beta,
        objective after the update).  trace: optional list receiving, per
        iteration, the normalised variables before the update, the gradients
        and the new direction."""
        vars = {k: np.array(v) for k, v in vars.items()}
        vars["obj"] = vars["obj"] / self.norm_const
        if start_iter == 0 and mask is not None:
            vars["obj"] = vars["obj"] * mask
        vars["proj"] = self.fwd_tomo(vars["obj"])
        etas = None
        hist = []
        for i in range(start_iter, niter):
            grads = self.gradients(vars)
            grads["proj"] = self.fwd_tomo(grads["obj"])
            if i == start_iter:
                beta = 0.0
            else:
                beta = self.hessian(vars, grads, etas) / self.hessian(vars, etas, etas)
            if etas is None:
                etas = {k: np.zeros_like(v) for k, v in grads.items()}
            top = 0.0
            for v in ["obj", "pos", "prb"]:
                etas[v] = beta * etas[v] - grads[v]
                top -= redot(grads[v], etas[v]) / self.rho_sq[v]
            etas["proj"] = beta * etas["proj"] - grads["proj"]
            bottom = self.hessian(vars, etas, etas)
            alpha = top / bottom
            if trace is not None:
                trace.append({"vars": {k: v.copy() for k, v in vars.items()},
                              "grads": grads, "etas": {k: v.copy() for k, v in etas.items()},
                              "alpha": alpha, "beta": beta})
            for v in ["obj", "prb", "pos", "proj"]:
                vars[v] = vars[v] + alpha * etas[v]
            hist.append((alpha, beta, self.min(vars)))
        vars["obj"] = vars["obj"] * self.norm_const
        return vars, hist
