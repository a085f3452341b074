"""GPU check of the coarsest-level dense solver (banded Cholesky on FP64
MFMA + banded solves -> explicit inverse) against numpy on random SPD
banded matrices, through the C ABI."""
import ctypes

import numpy as np
import pytest
import torch

import top3d_amd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,bw", [(40, 39), (64, 20), (65, 20), (135, 41), (300, 100),
                                  (517, 200), (1000, 999)])
def test_dense_spd_inverse(n, bw):
    print(f"[gpu dense] n={n} bw={bw}", flush=True)
    lib = top3d_amd.library()
    rng = np.random.default_rng(n)
    B = rng.standard_normal((n, n))
    M = B @ B.T / n + np.eye(n)
    mask = np.abs(np.subtract.outer(np.arange(n), np.arange(n))) <= bw
    M = M * mask
    M = 0.5 * (M + M.T) + n * np.eye(n) * 0.1
    A = torch.from_numpy(M.copy()).cuda()
    rc = lib.top3d_dense_spd_inverse(ctypes.c_void_p(A.data_ptr()), n, bw)
    assert rc == 0, top3d_amd.library().top3d_last_error_string()
    inv = A.cpu().numpy()
    ref = np.linalg.inv(M)
    assert np.abs(inv - ref).max() <= 1e-12 * np.abs(ref).max()
