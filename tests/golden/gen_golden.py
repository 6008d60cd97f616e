"""Writes the golden vectors under tests/golden/ from library routines in this
container (numpy, scipy) -- independent of the oracle and of the product:

  gauss_legendre.npz   numpy.polynomial.legendre.leggauss nodes/weights, n = 1..12
  jacobi.npz           scipy.special.eval_jacobi P_n^{(a,b)}(z) on a grid of z for
                       the (n, a, b) the Dubiner basis uses at degree <= 8
Run: python tests/golden/gen_golden.py
"""
import os

import numpy as np
from scipy.special import eval_jacobi

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    gl = {}
    for n in range(1, 13):
        x, w = np.polynomial.legendre.leggauss(n)
        gl[f"x{n}"] = x
        gl[f"w{n}"] = w
    np.savez(os.path.join(HERE, "gauss_legendre.npz"), **gl)

    z = np.linspace(-1.0, 1.0, 41)
    cases = []
    for n in range(0, 10):
        for a, b in [(0, 0), (1, 1), (2, 1), (2, 2)] + [(2 * i + 1, 0) for i in range(9)] + [(2 * i + 2, 1) for i in range(9)]:
            cases.append((n, a, b))
    cases = np.array(cases, dtype=np.int64)
    vals = np.array([eval_jacobi(n, a, b, z) for n, a, b in cases])
    np.savez(os.path.join(HERE, "jacobi.npz"), z=z, cases=cases, vals=vals)


if __name__ == "__main__":
    main()
