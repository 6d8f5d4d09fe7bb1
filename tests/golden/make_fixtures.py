"""Writes the golden fixtures of this directory.

The reference ships no golden vectors, and its own build needs an external
LAPACK/BLAS, so the fixtures are produced by an independent library routine:
LAPACK DGGHD3 (scipy's OpenBLAS), whose Hessenberg-triangular form equals the
reference's up to diagonal signs (both keep Q e1 = e1).  Each file holds the
seeded input pencil and the sign-normalised (H, T, Q, Z) with
T(j,j) >= 0 and H(j+1,j) >= 0, plus the (n, r, p, nq) to run the reference's
algorithm with.  Run:  python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from paraht_amd.pencils import illcond_pencil, random_pencil, uniform_pencil  # noqa: E402

CASES = [
    ("cfg0_small_n64_r16_p8_nq8", random_pencil, 64, 16, 8, 8),
    ("n40_r4_p3_nq3", random_pencil, 40, 4, 3, 3),
    ("n37_r5_p2_nq5", random_pencil, 37, 5, 2, 5),
    ("uniform_n48_r8_p4_nq4", uniform_pencil, 48, 8, 4, 4),
    ("illcond_n50_r4_p3_nq4", illcond_pencil, 50, 4, 3, 4),
]


def main():
    for name, gen, n, r, p, nq in CASES:
        A, B = gen(n, seed=1302)
        H, T, Q, Z = O.dgghd3(A, B)
        H, T, Q, Z = O.sign_normalize(H, T, Q, Z)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), A=A, B=B, H=H,
                            T=T, Q=Q, Z=Z, params=np.array([n, r, p, nq]))
        print("wrote", name)


if __name__ == "__main__":
    main()
