"""Writes the committed fixtures under tests/golden (run from the repo root).

eig_psd.npz: random symmetric 12x12 matrices (seeded), their eigenvalues from
numpy.linalg.eigh (LAPACK dsyevd) and their projection onto the PSD cone
V max(w, 0) V^T -- the quantity makePD (Math/UTILS.h:10-27) computes.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20240611)
    As, ws, Ps = [], [], []
    for k in range(24):
        M = rng.standard_normal((12, 12)) * (10.0 ** rng.uniform(-3, 3))
        A = M + M.T
        if k % 4 == 0:      # some nearly PSD matrices with tiny negative eigenvalues
            w, V = np.linalg.eigh(A)
            w = np.abs(w)
            w[0] = -1e-9 * w.max()
            A = (V * w) @ V.T
            A = 0.5 * (A + A.T)
        w, V = np.linalg.eigh(A)
        As.append(A)
        ws.append(w)
        Ps.append((V * np.maximum(w, 0)) @ V.T)
    np.savez(os.path.join(HERE, "eig_psd.npz"), A=np.array(As), w=np.array(ws), P=np.array(Ps))


if __name__ == "__main__":
    main()
