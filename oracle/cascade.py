"""Oracle: the BH cascade, its gradient and bilinear form (TEST INFRASTRUCTURE ONLY).

Restates src/holotomocupy/rec_mpi.py:517-813 (F0..F3, dF*, d2F_dF*, gF*),
rec_mpi.py:327-353 (gradients_cascade, the "BH gradient eval") and
rec_mpi.py:250-303 (hessian_cascade) for one chunk of angles, on the oracle
operators (oracle.propagation, oracle.shift):

    F3: x22[:, j] = S(coeff(proj), pos[:, j], m[:, j])
    F2: x12 = exp(i x22)
    F1: x0[:, j] = D_j(prb_j * x12[:, j])
    F0: sum((|x0| - d)^2) / data_size

Variables are x = [prb (ndist, nz, n), proj (B, nzobj, nobj), pos (B, ndist, 2)],
data d = sqrt(intensity) (B, ndist, nz, n).  ``m`` is the per-angle
magnification of each distance (eff_demagnifications, rec_mpi.py:146), given
here per chunk angle.

Pins (tests/test_oracle_cascade.py):
  * the gradient matches central finite differences of the objective along
    random directions (all three variables);
  * the bilinear form with equal directions matches second differences;
  * Taylor ratios of each stage (tests/holotomo3d/test_approximation.ipynb):
    first-order remainders scale by 4 and second-order ones by 8 as h halves;
  * the bilinear form is symmetric in its two directions.
"""
import numpy as np


def reprod(a, b):
    """utils.py:96-97"""
    return np.real(a) * np.real(b) + np.imag(a) * np.imag(b)


def redot(a, b):
    """utils.py:100-103 (axis=None)"""
    return float(np.sum(reprod(a, b)))


class Cascade:
    """Per-chunk cascade of rec_mpi.Rec with the oracle operators."""

    def __init__(self, prop, shift, ndist, nz, n, nzobj, nobj, data_size, mag, rho=(1.0, 1.0, 1.0),
                 obj_real=False, single=False):
        """single: complex64 intermediates (with a complex64 propagator and a
        single-precision shift), for timing the reference's precision on the
        CPU; the default is FP64 throughout."""
        self.prop, self.shift = prop, shift
        self.cdtype = np.complex64 if single else np.complex128
        self.rdtype = np.float32 if single else np.float64
        self.ndist, self.nz, self.n, self.nzobj, self.nobj = ndist, nz, n, nzobj, nobj
        self.data_size = data_size
        self.mag = np.asarray(mag, dtype=np.float64)  # [B, ndist]
        self.rho_sq = {"obj": rho[0] ** 2, "prb": rho[1] ** 2, "pos": rho[2] ** 2}
        self.obj_real = obj_real
        self.F = [self.F0, self.F1, self.F2, self.F3]
        self.gF = [self.gF0, self.gF1, self.gF2, self.gF3]
        self.dF = [self.dF0, self.dF1, self.dF2, self.dF3]
        self.d2F_dF = [self.d2F_dF0, self.d2F_dF1, self.d2F_dF2, self.d2F_dF3]

    # ---- F0 = 1/N || |x0| - d ||^2   (rec_mpi.py:517-569)
    def F0(self, x, d):
        return float(np.sum((np.abs(x) - d) ** 2) / self.data_size)

    def dF0(self, x, y, d, return_x=False):
        return 2 / self.data_size * redot(x - d * (x / np.abs(x)), y)

    def d2F_dF0(self, x, y, z, w, d):
        absval = np.abs(x)
        l0 = x / absval
        d0 = d / absval
        v = (1 - d0) * reprod(y, z) + d0 * reprod(l0, y) * reprod(l0, z)
        if w is not None:
            v = v + reprod(x - d * l0, w)
        return 2 / self.data_size * float(np.sum(v))

    def gF0(self, x, y):
        for i in range(1, 4)[::-1]:
            x = self.F[i](x)
        return 2 / self.data_size * (x - y * (x / np.abs(x)))

    # ---- x0 = F1(x11, x12) = D(x11 x12)   (rec_mpi.py:572-643)
    def F1(self, x):
        x11, x12 = x
        x0 = np.empty([x12.shape[0], self.ndist, self.nz, self.n], dtype=self.cdtype)
        for j in range(self.ndist):
            x0[:, j] = self.prop.D(x11[j] * x12[:, j], j)
        return x0

    def dF1(self, x, y, return_x=True):
        x11, x12 = x
        y11, y12 = y
        y0 = y11[None] * x12 + x11[None] * y12
        for j in range(self.ndist):
            y0[:, j] = self.prop.D(y0[:, j], j)
        if return_x:
            x0 = x11[None] * x12
            for j in range(self.ndist):
                x0[:, j] = self.prop.D(x0[:, j], j)
            return x0, y0
        return y0

    def d2F_dF1(self, x, y, z, w):
        x11, x12 = x
        y11, y12 = y
        z11, z12 = z
        w11, w12 = w
        if y12 is z12:
            y0 = 2 * y11[None] * y12
        else:
            y0 = y11[None] * z12 + z11[None] * y12
        if w11 is not None:
            y0 = y0 + w11[None] * x12
        if w12 is not None:
            y0 = y0 + x11[None] * w12
        for j in range(self.ndist):
            y0[:, j] = self.prop.D(y0[:, j], j)
        return y0

    def gF1(self, x, y):
        y0 = y
        for i in range(2, 4)[::-1]:
            x = self.F[i](x)
        x11, x12 = x
        y12 = np.empty([y0.shape[0], self.ndist, self.nz, self.n], dtype=self.cdtype)
        for j in range(self.ndist):
            y12[:, j] = self.prop.DT(y0[:, j], j)
        y11 = np.sum(y12 * np.conj(x12), axis=0)
        y12 = y12 * np.conj(x11[None])
        return y11, y12

    # ---- (x11, x12) = F2(x21, x22) = (x21, e^{i x22})   (rec_mpi.py:645-724)
    def F2(self, x):
        x21, x22 = x
        return x21, np.exp(1j * x22)

    def dF2(self, x, y, return_x=True):
        x21, x22 = x
        y21, y22 = y
        x12 = np.exp(1j * x22)
        y12 = x12 * 1j * y22
        return ([x21, x12], [y21, y12]) if return_x else [y21, y12]

    def d2F_dF2(self, x, y, z, w):
        x21, x22 = x
        y21, y22 = y
        z21, z22 = z
        w21, w22 = w
        y12 = np.exp(1j * x22) * (-y22 * z22)
        if w22 is not None:
            y12 = y12 + np.exp(1j * x22) * 1j * w22
        return [w21, y12]

    def gF2(self, x, y):
        y11, y12 = y
        x = self.F[3](x)
        x21, x22 = x
        y22 = (-1j) * y12 * np.conj(np.exp(1j * x22))
        if self.obj_real:
            y22 = y22.real
        return [y11, y22]

    # ---- (x21, x22) = F3(x31, x32, x33) = (x31, S_x33(x32))   (rec_mpi.py:726-813)
    def F3(self, x):
        x31, x32, x33 = x
        x22 = np.empty([len(x33), self.ndist, self.nz, self.n],
                       dtype=self.rdtype if self.obj_real else self.cdtype)
        c = self.shift.coeff(x32)
        for k in range(self.ndist):
            x22[:, k] = self.shift.curlySc(c, x33[:, k], self.mag[:, k])
        return [x31, x22]

    def dF3(self, x, y, return_x=True):
        x31, x32, x33 = x
        y31, y32, y33 = y
        c = self.shift.coeff(x32)
        c1 = self.shift.coeff(y32)
        y22 = np.stack([self.shift.dcurlySc(c, x33[:, k], self.mag[:, k], c1, y33[:, k])
                        for k in range(self.ndist)], axis=1)
        if return_x:
            x22 = np.stack([self.shift.curlySc(c, x33[:, k], self.mag[:, k]) for k in range(self.ndist)], axis=1)
            return [x31, x22], [y31, y22]
        return [y31, y22]

    def d2F_dF3(self, x, y, z, w):
        x31, x32, x33 = x
        y31, y32, y33 = y
        z31, z32, z33 = z
        w31, w32, w33 = w
        c = self.shift.coeff(x32)
        cy = self.shift.coeff(y32)
        cz = self.shift.coeff(z32)
        y22 = np.stack([self.shift.d2curlySc(c, x33[:, k], self.mag[:, k], cy, y33[:, k], cz, z33[:, k])
                        for k in range(self.ndist)], axis=1)
        if w32 is not None:
            cw = self.shift.coeff(w32)
            y22 = y22 + np.stack([self.shift.dcurlySc(c, x33[:, k], self.mag[:, k], cw, w33[:, k])
                                  for k in range(self.ndist)], axis=1)
        return [w31, y22]

    def gF3(self, x, y):
        y21, y22 = y
        x31, x32, x33 = x
        y32 = np.zeros([y22.shape[0], self.nzobj, self.nobj],
                       dtype=self.rdtype if self.obj_real else self.cdtype)
        y33 = np.empty([y22.shape[0], self.ndist, 2])
        c = self.shift.coeff(x32)
        for k in range(self.ndist):
            dpsi, dr = self.shift.dcurlySadjc(c, x33[:, k], self.mag[:, k], y22[:, k])
            y32 += dpsi
            y33[:, k] = dr
        y32 = self.shift.coeff(y32)
        return [y21, y32, y33]

    # ---- drivers
    def objective(self, prb, proj, pos, d):
        """Data term of Rec.min (rec_mpi.py:815-835) for this chunk."""
        y = [prb, proj, pos]
        for i in range(1, 4)[::-1]:
            y = self.F[i](y)
        return self.F0(y, d)

    def forward(self, prb, proj, pos):
        """F1(F2(F3(x))), whose modulus is the synthetic data (rec_mpi.py:882-896)."""
        y = [prb, proj, pos]
        for i in range(1, 4)[::-1]:
            y = self.F[i](y)
        return y

    def gradients(self, prb, proj, pos, d):
        """One BH gradient evaluation of the chunk (rec_mpi.py:327-353):
        returns (gradprb, gradproj, gradpos), each scaled by rho^2."""
        x = [prb, proj, pos]
        y = d
        for i in range(4):
            y = self.gF[i](x, y)
        return (y[0] * self.rho_sq["prb"], y[1] * self.rho_sq["obj"], y[2] * self.rho_sq["pos"])

    def hessian(self, x, y, z, d):
        """BH bilinear form of the data term (rec_mpi.py:250-303); x, y, z are
        [prb, proj, pos] triples, y is z when the caller passes the same list."""
        y_is_z = y is z
        w = [None, None, None]
        x, y, z = list(x), list(y), list(z) if not y_is_z else None
        if y_is_z:
            z = y
        for i in range(1, 4)[::-1]:
            w = self.d2F_dF[i](x, y, z, w)
            fx, y = self.dF[i](x, y)
            z = y if y_is_z else self.dF[i](x, z, return_x=False)
            x = fx
        return self.d2F_dF0(x, y, z, w, d)
