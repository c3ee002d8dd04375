'''Rotational (Raman) nonlinear susceptibility of linear molecules (modules/rotations.py:13-85).

A ``Rotator`` holds the level tables of the density-matrix model; the stepper evaluates the susceptibility it
induces on every time strip (kernels/rotations.cl:80-129, AddChiUPPE) inside the source term. The tables are the
reference's:

* rotational frequencies w_j = 2 pi B j (j + 1), j = 0 .. n-1 (rotations.py:27);
* thermal populations rho0_j = d_j exp(-hbar w_j / kT) / Z, Z = sum (2 j + 1) d_j exp(-hbar w_j / kT)
  (rotations.py:28-30), zero when Z = 0;
* for the j <-> j-2 transitions (j >= 2): W_j = w_j - w_{j-2} + i damp_j and T_j = rho0_j - rho0_{j-2}
  (rotations.py:33-35).

The constructor takes B, the polarizability difference, kT, the nuclear degeneracy array (its length sets the
state count), the per-level damping and the normalised hbar, in the reference's argument order.
'''
import numpy as np


class Rotator:
    def __init__(self, B, dalpha, kT, dnuc, damp, hbar):
        dnuc = np.asarray(dnuc, dtype=np.float64)
        damp = np.asarray(damp, dtype=np.float64)
        n = dnuc.shape[0]
        j = np.arange(0, n, 1)
        self.dalpha = float(dalpha)
        self.hbar = float(hbar)
        self.wj = 2 * np.pi * B * j * (j + 1)
        X = dnuc * np.exp(-self.hbar * self.wj / kT)
        Z = np.sum((2 * j + 1) * X)
        self.rho0jm = X / Z if Z > 0 else np.zeros(n)
        self.Wj = np.zeros(n, dtype=np.complex128)
        self.Tj = np.zeros(n)
        for k in range(2, n):
            self.Wj[k] = self.wj[k] - self.wj[k - 2] + 1j * damp[k]
            self.Tj[k] = self.rho0jm[k] - self.rho0jm[k - 2]

    @property
    def states(self):
        return self.Wj.shape[0]
