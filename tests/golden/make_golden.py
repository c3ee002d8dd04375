'''Writes the golden fixtures of this folder: outputs of the FP64 oracle on small seeded workloads.

Run from the repository root after building the oracle (python oracle/build.py):
    python tests/golden/make_golden.py

Each .npz holds, for the last diagnostic plane of the case, the r-space field A and the listed diagnostics
(complex128, the reference layout (Nw, N1, N2); the paraxial cases are envelopes, frequencies sorted), the dz of every step taken (record), and the RHS evaluation
count per plane. The fixtures freeze the oracle's results: the CPU tests check that the oracle still
reproduces them, the GPU tests compare the HIP library with them.
'''
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from searay_amd import configs  # noqa: E402

CASES = {
    # radial Kerr in air, Nw = 65, Nr = 16, 4 fixed steps per plane over 1 m: A, J, chiNL, ne
    'radial_kerr_65x16': (lambda: configs.radial_kerr_air(scale=8, steps=4), ('J', 'chi', 'ne')),
    # radial Kerr + O2 photoionisation (closed-form rate), Nw = 513, Nr = 16, 3 subcycles over 1 mm with
    # adaptive steps (10 steps): A, ne
    'radial_o2_explicit_513x16': (lambda: configs.radial_o2_ionization(scale=4, steps=3, length='1 mm', nr=16),
                                  ('ne',)),
    # the same with the series (quintic-fit) rate: A
    'radial_o2_series_513x16': (lambda: configs.radial_o2_ionization(scale=4, steps=3, length='1 mm', nr=16,
                                                                     rate='series'), ()),
    # radial Kerr in N2 with the rotational (Raman) susceptibility of 32 states, Nw = 65, Nr = 16: A, chiNL
    'radial_raman_65x16': (lambda: configs.radial_raman_air(scale=8, steps=4), ('chi',)),
    # 3-D Cartesian filamentation (super-Gaussian with amplitude noise, Kerr + O2), Nw = 33, 16 x 16, 3 fixed
    # steps over 1 mm: A
    'cartesian_fil_33x16x16': (lambda: configs.cartesian_filamentation(scale=16, steps=3, length='1 mm'), ()),
    # paraxial envelope: the focused air filament (Kerr + cycle-averaged O2 rate), Nw = 64, Nr = 64 with 32
    # modes, 50 mJ, one plane of 4 mm with adaptive steps: A, J, chiNL, ne
    'paraxial_fil_64x64': (lambda: configs.paraxial_air_filament(scale=4, planes=2, length='4 mm', power=50.0),
                           ('J', 'chi', 'ne')),
    # the same with the 32-state N2 rotator: A, chiNL
    'paraxial_raman_64x64': (lambda: configs.paraxial_air_filament(scale=4, planes=2, length='4 mm', power=50.0,
                                                                   rotator_states=32), ('chi',)),
    # paraxial envelope in 3-D Cartesian, Nw = 16, 16 x 16, 200 mJ: A, ne
    'paraxial_cart_16x16x16': (lambda: configs.paraxial_air_filament(scale=16, planes=2, length='4 mm', power=200.0,
                                                                     coords='cartesian'), ('ne',)),
}


def main():
    from oracle import oracle
    for name, (make, diag) in CASES.items():
        rec = []
        A, J, chi, ne, dom4d, evals = oracle.oracle_track(make(), record=rec)
        path = os.path.join(HERE, name + '.npz')
        out = dict(A=A[..., -1], dz=np.array(rec, dtype=np.float64), evals=np.array(evals, dtype=np.int64))
        full = dict(J=J, chi=chi, ne=ne)
        for d in diag:
            out[d] = full[d][..., -1]
        np.savez_compressed(path, **out)
        print('%s: %d steps, %d bytes' % (name, len(rec), os.path.getsize(path)))


if __name__ == '__main__':
    main()
