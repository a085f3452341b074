"""CPU check of the level-1 coefficient tables compiled into the fine-tile
kernels (topopt-in-openmp-gpu_amd/csrc/k_fine_tile.hip): kK1Mask, the
non-zero pattern of the transformed level-1 sub-cell matrix K1, and
kK1Class, the magnitude class and sign of each non-zero.

K1 is the Walsh-Hadamard transform of sub-cell 0 of KE_1 (getKEsubspace,
local_matrix.c:251-319), with rows and columns relabelled sigma = q ^ (1 << c).
The pattern and the classes must hold for every Poisson ratio: two non-zeros
share a class when they are equal up to sign for all ratios sampled here, so
a table derived from one ratio only, or a transposed or shifted table, fails.
No GPU call is made."""
import ctypes
import os
import re

import numpy as np

import top3d_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "topopt-in-openmp-gpu_amd", "csrc", "k_fine_tile.hip")
POS = [(0, 1, 0), (1, 1, 0), (1, 0, 0), (0, 0, 0), (0, 1, 1), (1, 1, 1), (1, 0, 1), (0, 0, 1)]
CHI = np.array([[-1.0 if (sum(((q >> b) & 1) * POS[n][b] for b in range(3)) & 1) else 1.0
                 for n in range(8)] for q in range(8)])
NUS = [0.05, 0.1, 0.17, 0.2, 0.25, 0.3, 0.33, 0.35, 0.4, 0.45, 0.49]


def k1(nu):
    ke = np.zeros(576 * 8)
    top3d_amd.library().getKEsubspace(ke.ctypes.data_as(ctypes.c_void_p), nu, 1)
    K = ke[:576].reshape(8, 3, 8, 3)  # sub-cell 0
    out = np.zeros((24, 24))
    for sa in range(8):
        for ca in range(3):
            for sb in range(8):
                for cb in range(3):
                    qa, qb = sa ^ (1 << ca), sb ^ (1 << cb)
                    out[3 * sa + ca, 3 * sb + cb] = CHI[qa] @ K[:, ca, :, cb] @ CHI[qb] / 64
    return out


def compiled_tables():
    text = open(SRC).read()
    mask = re.search(r"kK1Mask\[24\]\s*=\s*\{(.*?)\};", text, re.S).group(1)
    mask = [int(v, 16) for v in re.findall(r"0x[0-9a-fA-F]+", mask)]
    cls = re.search(r"kK1Class\[kK1Nnz\]\s*=\s*\{(.*?)\};", text, re.S).group(1)
    cls = [int(v) for v in re.findall(r"\d+", cls)]
    ncls = int(re.search(r"constexpr int kK1Classes = (\d+);", text).group(1))
    return mask, cls, ncls


def derived_tables():
    Ks = [k1(nu) for nu in NUS]
    nz = np.abs(Ks[5]) > 1e-12 * np.abs(Ks[5]).max()
    for K in Ks:  # the pattern does not depend on the ratio
        assert np.array_equal(np.abs(K) > 1e-12 * np.abs(K).max(), nz)
    mask = [sum(1 << c for c in range(24) if nz[r, c]) for r in range(24)]
    V = np.array([K[nz] for K in Ks])  # row-major order of the non-zeros
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
    return mask, code, len(reps)


def test_k1_tables_hold_for_every_poisson_ratio():
    mask, cls, ncls = compiled_tables()
    dmask, dcls, dn = derived_tables()
    assert mask == dmask
    assert sum(bin(m).count("1") for m in mask) == len(cls) == 279
    assert ncls == dn == 20
    assert cls == dcls
