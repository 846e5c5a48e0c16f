"""Writes tests/golden/linalg_golden.npz: reference values for the 6x6
solve and the SE(3) exponential, computed with numpy / scipy (independent
libraries, not the reference project).

    python tests/golden/gen_golden.py
"""
import os

import numpy as np
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))


def hat(xi):
    rho, w = xi[:3], xi[3:]
    M = np.zeros((4, 4))
    M[:3, :3] = [[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]]
    M[:3, 3] = rho
    return M


def main():
    rng = np.random.default_rng(20261202)
    # symmetric positive definite systems with a wide range of conditioning
    A, b, x = [], [], []
    for i in range(64):
        J = rng.normal(size=(40, 6)) * np.logspace(0, rng.uniform(0, 4), 6)
        H = J.T @ J
        r = rng.normal(size=6)
        A.append(H)
        b.append(r)
        x.append(np.linalg.solve(H, r))
    # twists: tiny, small and large rotations
    xis, Ts = [], []
    for scale in (1e-12, 1e-9, 1e-6, 1e-3, 0.1, 1.0, 3.0):
        for _ in range(8):
            xi = rng.normal(size=6) * scale
            xis.append(xi)
            Ts.append(scipy.linalg.expm(hat(xi)))
    np.savez(os.path.join(HERE, "linalg_golden.npz"), A=np.array(A), b=np.array(b), x=np.array(x),
             xi=np.array(xis), T=np.array(Ts))


if __name__ == "__main__":
    main()
