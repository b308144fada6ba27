"""Writes tests/golden/near_field_N.npz and near_field_K.npz: the near-field
N- and K-operator entries of the reference's surface-surface formulation at
the offsets {0,1}^3, computed by the oracle's restatement
(oracle/near_field.py) around the reference's own DIRECTFN C++ (compiled from
the reference checkout by oracle/Makefile into oracle/_ref/).  Run in the
container that holds the reference:

    make -C oracle && python tests/golden/make_near_field.py [N] [K]

Cases: 5 mm voxels at 298 MHz (the 32^3 sphere workload) for PWC and PWL,
4 mm at 128 MHz (the 64x64x80 head workload) for PWL; Np_1D_near_S = 15
(load_inputs.m:80).  Stored per case: res, k0, Np, pwl and the entries
A[m1, m2, m3, pair, component] (pairs in the order of VV_Nop.m:27-28; N
components xx xy xz yy yz zz, K components the (p, q) of
surface_surface_coeff_Kop.m:10-11).
"""
import itertools
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import near_field as nf  # noqa: E402
from oracle import vie_oracle as vo  # noqa: E402

CASES = [("c1_pwc", 0.005, 298e6, False), ("c1_pwl", 0.005, 298e6, True), ("c2_pwl", 0.004, 128e6, True)]
NP = 15


def main(ops):
    lib = nf._lib()
    for op in ops:
        near = nf.near_N if op == "N" else nf.near_K
        out = {}
        for name, res, freq, pwl in CASES:
            k0 = vo.em_constants(freq)["k0"]
            A = np.zeros((2, 2, 2, 10 if pwl else 1, 6 if op == "N" else 3), dtype=np.complex128)
            t0 = time.time()
            for m in itertools.product((0, 1), repeat=3):
                A[m] = near(tuple(np.array(m) + 1), res, k0, 4 if pwl else 1, NP, lib)
            print(f"{op} {name}: {time.time() - t0:.0f} s", flush=True)
            out[name + "_A"] = A
            out[name + "_params"] = np.array([res, k0, NP, float(pwl)])
        np.savez_compressed(os.path.join(HERE, f"near_field_{op}.npz"), **out)


if __name__ == "__main__":
    main(sys.argv[1:] or ["N", "K"])
