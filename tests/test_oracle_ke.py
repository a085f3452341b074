"""Pins the oracle's element matrices (or_ke, restating local_matrix.c:251-319)
against an independent tensor rule, symmetry and the rigid-body null space."""
import itertools

import numpy as np
import pytest

import fem_ref
import pyoracle


@pytest.fixture(scope="module")
def grid():
    g = pyoracle.Grid(8, 4, 4, 3)
    yield g
    g.close()


@pytest.mark.parametrize("l", [0, 1, 2])
def test_subcell_matrices_match_independent_rule(grid, l):
    ref = fem_ref.subcell_matrices(l)
    got = grid.ke(l) if l < 3 else None
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-13 * scale


def test_other_point_counts_match_moment_rule():
    # the midpoint rule with n points equals the moment-matched 2-point rule
    # for every n, not only n = 10: a wrong quadrature spacing fails here
    lib = pyoracle.load("fp64")
    for npts in (3, 7):
        out = np.zeros(576)
        lib.or_ke(0.3, 0, npts, out.ctypes.data_as(pyoracle.c_void_p))
        ref = fem_ref.element_matrix(1.0, [(0, 1)] * 3, npts)
        assert np.abs(out.reshape(24, 24) - ref).max() <= 1e-13 * np.abs(ref).max()


def test_ke_symmetric_psd_six_rigid_modes(grid):
    K = grid.ke(0)[0]
    assert np.abs(K - K.T).max() <= 1e-15 * np.abs(K).max()
    w = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert (w > -1e-14).all()
    assert (np.abs(w) < 1e-13).sum() == 6
    # rigid modes in the reference's convention (y displacement along -y)
    pos = fem_ref.LNODE.astype(float)
    modes = []
    for d in range(3):
        m = np.zeros((8, 3))
        m[:, d] = 1
        modes.append(m)
    for a, b in ((0, 1), (1, 2), (0, 2)):
        m = np.zeros((8, 3))
        m[:, a], m[:, b] = -pos[:, b], pos[:, a]
        modes.append(m)
    for m in modes:
        v = (m * [1, -1, 1]).ravel()
        assert np.abs(K @ v).max() <= 1e-13


def test_subcell_sum_is_galerkin_of_fine_matrix(grid):
    # KE_1[cell] = Q_cell^T KE_0 Q_cell, Q = trilinear interpolation from the
    # parent's nodes to the child's nodes (same quadrature points)
    K0 = grid.ke(0)[0]
    K1 = grid.ke(1)
    pos = fem_ref.LNODE
    for ii, kk, jj in itertools.product(range(2), repeat=3):
        Q = np.zeros((24, 24))
        for a in range(8):  # child node
            xc = (np.array([ii, jj, kk]) + pos[a]) / 2.0
            for b in range(8):  # parent node
                w = np.prod([xc[d] if pos[b][d] else 1 - xc[d] for d in range(3)])
                for c in range(3):
                    Q[3 * a + c, 3 * b + c] = w
        ref = Q.T @ K0 @ Q
        got = K1[4 * ii + 2 * kk + jj]
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
