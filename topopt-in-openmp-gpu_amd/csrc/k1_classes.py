"""Prints kK1Class for k_fine_tile.hip: for each non-zero of the transformed
level-1 sub-cell matrix K1 (row-major over the kK1Mask pattern), the index of
its magnitude class and its sign relative to the class's first member.

Two non-zeros share a class when they are equal up to sign for every Poisson
ratio sampled below, so the classes come from the symmetry of the cube, not
from one ratio.  The library checks the table against the K1 it computes at
start-up and refuses the level-1 kernel if it does not hold.

Usage: python k1_classes.py   (needs the built lib/libtop3d_hip.so; no GPU)"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = ctypes.CDLL(os.path.join(HERE, "..", "lib", "libtop3d_hip.so"))
LIB.getKEsubspace.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
LIB.getKEsubspace.restype = None
# local node positions (x, y, z) of the element's 8 nodes
POS = [(0, 1, 0), (1, 1, 0), (1, 0, 0), (0, 0, 0), (0, 1, 1), (1, 1, 1), (1, 0, 1), (0, 0, 1)]
CHI = np.array([[-1.0 if (sum(((q >> b) & 1) * POS[n][b] for b in range(3)) & 1) else 1.0
                 for n in range(8)] for q in range(8)])


def k1(nu):
    ke = np.zeros(576 * 8)
    LIB.getKEsubspace(ke.ctypes.data_as(ctypes.c_void_p), nu, 1)
    K = ke[:576].reshape(8, 3, 8, 3)  # sub-cell 0
    out = np.zeros((24, 24))
    for sa in range(8):
        for ca in range(3):
            for sb in range(8):
                for cb in range(3):
                    qa, qb = sa ^ (1 << ca), sb ^ (1 << cb)
                    out[3 * sa + ca, 3 * sb + cb] = CHI[qa] @ K[:, ca, :, cb] @ CHI[qb] / 64
    return out


def main():
    nus = [0.05, 0.1, 0.17, 0.2, 0.25, 0.3, 0.33, 0.35, 0.4, 0.45, 0.49]
    Ks = [k1(nu) for nu in nus]
    mask = np.abs(Ks[5]) > 1e-12 * np.abs(Ks[5]).max()
    V = np.array([K[mask] for K in Ks])  # row-major order of the non-zeros
    tol = 1e-10 * np.abs(V).max(axis=1)
    reps, code = [], []
    for t in range(V.shape[1]):
        for c, r in enumerate(reps):
            s = next((s for s in (1, -1) if np.all(np.abs(V[:, t] - s * V[:, r]) <= tol)), 0)
            if s:
                code.append(c | (128 if s < 0 else 0))
                break
        else:
            reps.append(t)
            code.append(len(reps) - 1)
    print(f"// {len(code)} non-zeros in {len(reps)} classes")
    print(f"constexpr int kK1Classes = {len(reps)};")
    print("static constexpr uint8_t kK1Class[kK1Nnz] = {")
    for a in range(0, len(code), 16):
        print("    " + ", ".join(str(v) for v in code[a:a + 16]) + ",")
    print("};")


if __name__ == "__main__":
    main()
