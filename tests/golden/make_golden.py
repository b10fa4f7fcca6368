"""Writes tests/golden/golden.npz: library-computed golden vectors.

* Wigner 3j symbols from sympy's exact rational arithmetic
  (sympy.physics.wigner.wigner_3j), l up to 60.
* Spherical Bessel j_l, y_l from scipy.special.spherical_jn / spherical_yn for
  l <= 60 over the argument range the solver meets (k r from 5e-4 to 200).

These are third-party library values (not the reference's code); they pin the
oracle's and the product's own implementations of the published formulas.
Run: python tests/golden/make_golden.py
"""
import os
import random

import numpy as np
from scipy.special import spherical_jn, spherical_yn
from sympy.physics.wigner import wigner_3j


def main():
    rng = random.Random(20260)
    w3 = []
    while len(w3) < 600:
        j1, j2 = rng.randint(0, 60), rng.randint(0, 60)
        j3 = rng.randint(abs(j1 - j2), min(j1 + j2, 120))
        m1, m2 = rng.randint(-j1, j1), rng.randint(-j2, j2)
        m3 = -m1 - m2
        if abs(m3) > j3:
            continue
        w3.append((j1, j2, j3, m1, m2, m3, float(wigner_3j(j1, j2, j3, m1, m2, m3))))
    # all (l1 l2 l3; 0 0 0) and a dense low-l block
    for j1 in range(0, 9):
        for j2 in range(0, 9):
            for j3 in range(abs(j1 - j2), j1 + j2 + 1):
                for m1 in range(-j1, j1 + 1):
                    for m2 in range(-j2, j2 + 1):
                        m3 = -m1 - m2
                        if abs(m3) <= j3 and rng.random() < 0.05:
                            w3.append((j1, j2, j3, m1, m2, m3, float(wigner_3j(j1, j2, j3, m1, m2, m3))))
    xs = np.array([5e-4, 1e-3, 0.01, 0.05, 0.1, 0.3, 0.5, 1.0, 2.0, 2.7, 5.0, 7.5, 10.0, 15.0, 29.9, 30.5,
                   45.0, 60.0, 90.0, 100.0, 120.0, 150.0, 200.0])
    ls = np.arange(0, 61)
    J = np.array([[x * spherical_jn(int(l), x) for x in xs] for l in ls])
    with np.errstate(over="ignore"):
        Y = np.array([[x * spherical_yn(int(l), x) for x in xs] for l in ls])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")
    np.savez_compressed(out, w3j=np.array(w3), bessel_x=xs, bessel_l=ls, ric_j=J, ric_y=Y)
    print("wrote", out, len(w3), "3j values")


if __name__ == "__main__":
    main()
