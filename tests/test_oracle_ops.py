"""Pins the oracle's operators (stencil_methods.c, stencil_assembly.c)
against the independent scipy model in fem_ref."""
import numpy as np
import pytest
import scipy.sparse as sp

import fem_ref
import pyoracle

NELX, NELY, NELZ, NL = 8, 4, 4, 3


def rand_density(seed, halo=0.7):
    rng = np.random.default_rng(seed)
    vals = rng.uniform(0.1, 1.0, NELX * NELY * NELZ)
    return fem_ref.interior_density(NELX, NELY, NELZ, vals, halo=halo)


def rand_field(lay, seed):
    rng = np.random.default_rng(seed)
    v = np.zeros(lay.ndof)
    d = lay.interior_dofs()
    v[d] = rng.standard_normal(d.size)
    return v


@pytest.fixture(scope="module")
def grid():
    g = pyoracle.Grid(NELX, NELY, NELZ, NL)
    yield g
    g.close()


def test_fixed_dofs_and_layout(grid):
    for l in range(NL):
        lay = fem_ref.Layout(NELX, NELY, NELZ, l)
        assert grid.padding(l)[1:] == (lay.wx, lay.wy, lay.wz, lay.ndof)
        assert np.array_equal(np.sort(grid.fixed(l)), lay.fixed_dofs())
    # the reference's 128-wide y padding: wrapy = 128*m + 2
    g128 = pyoracle.Grid(NELX, NELY, NELZ, NL, yalign=128)
    assert g128.padding(0)[2] == 130
    g128.close()


def test_load():
    # too small a grid gives an empty load disc (as in the reference), so
    # use 16 x 8 x 8: radius (nely+1)/5 = 1.8 around (nelx, (nely-1)/2)
    g = pyoracle.Grid(16, 8, 8, 3)
    F = g.load()
    g.close()
    lay = fem_ref.Layout(16, 8, 8, 0)
    assert abs(F.sum() + 1.0) < 1e-14
    nz = np.nonzero(F)[0]
    assert (nz % 3 == 2).all()  # z direction only
    nodes = nz // 3
    ks = (nodes % (lay.wy * lay.wz)) // lay.wy
    iis = nodes // (lay.wy * lay.wz)
    assert (ks == 1).all()  # z = 0 face
    assert iis.min() >= 16 - 1  # near the x = nelx end
    # elements with (i - nelx)^2 + (j - 3.5)^2 < 1.8^2: i = 15, j = 2..5
    # (i - 16)^2 + (j - 3.5)^2 < 3.24 holds for i = 15, j = 3, 4 only
    # (j = 2, 5 give 3.25): 2 elements, nodal force -1/8, 6 nodes
    assert len(nz) == 6
    assert sorted(np.round(F[nz] * 8).astype(int)) == [-2, -2, -1, -1, -1, -1]


@pytest.mark.parametrize("seed", [1, 2])
def test_matvec0_matches_scipy(grid, seed):
    x = rand_density(seed)
    lay = fem_ref.Layout(NELX, NELY, NELZ, 0)
    K = fem_ref.assemble(NELX, NELY, NELZ, 0, x)
    Kbc = fem_ref.bc_identity_rows(K, lay.fixed_dofs())
    v = rand_field(lay, seed + 10)
    got = grid.matvec0(x, v)
    ref = Kbc @ v
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("l", [1, 2])
def test_level_operator_is_galerkin(grid, l):
    # K_l = R K_{l-1} P before BCs; matrix-free level operators keep identity
    # rows at fixed DOFs (stencil_methods.c:515-518)
    x = rand_density(3)
    lay = fem_ref.Layout(NELX, NELY, NELZ, l)
    Kf = fem_ref.assemble(NELX, NELY, NELZ, l - 1, x)
    P = fem_ref.prolongation(NELX, NELY, NELZ, l - 1)
    Kg = (P.T @ Kf @ P).tocsr()
    v = rand_field(lay, 11)
    got = grid.matvec_l(l, x, v)
    ref = Kg @ v
    fixed = lay.fixed_dofs()
    ref[fixed] = v[fixed]
    inter = lay.interior_dofs()
    assert np.abs(got[inter] - ref[inter]).max() <= 1e-12 * np.abs(ref).max()
    other = np.setdiff1d(np.arange(lay.ndof), inter)
    assert np.abs(got[other]).max() == 0.0


@pytest.mark.parametrize("l", [0, 1])
def test_transfers(grid, l):
    P = fem_ref.prolongation(NELX, NELY, NELZ, l)
    lc = fem_ref.Layout(NELX, NELY, NELZ, l + 1)
    lf = fem_ref.Layout(NELX, NELY, NELZ, l)
    vc = rand_field(lc, 5)
    assert np.abs(grid.prolong(l, vc) - P @ vc).max() <= 1e-15
    vf = rand_field(lf, 6)
    ref = P.T @ vf
    got = grid.restrict(l, vf)
    assert np.abs(got - ref).max() <= 1e-14
    # P reproduces a linear field exactly
    lin = np.zeros(lc.ndof)
    for (i, j, k) in lc.interior_nodes():
        lin[3 * lc.node(i, j, k)] = 2 * (i - 1) + 3 * 2 * (j - 1) - 2 * (k - 1)
    fine = grid.prolong(l, lin)
    for (i, j, k) in lf.interior_nodes():
        assert fine[3 * lf.node(i, j, k)] == pytest.approx((i - 1) + 3 * (j - 1) - (k - 1))


@pytest.mark.parametrize("l", [0, 1, 2])
def test_inverse_diagonal(grid, l):
    x = rand_density(7)
    lay = fem_ref.Layout(NELX, NELY, NELZ, l)
    K = fem_ref.assemble(NELX, NELY, NELZ, l, x)
    d = K.diagonal()
    d[lay.fixed_dofs()] = 1.0
    inter = lay.interior_dofs()
    got = grid.invdiag(x, l)
    assert np.abs(got[inter] * d[inter] - 1).max() <= 1e-13
    other = np.setdiff1d(np.arange(lay.ndof), inter)
    assert np.abs(got[other]).max() == 0.0


def test_csr_symmetric_elimination(grid):
    l = 2
    x = rand_density(8)
    lay = fem_ref.Layout(NELX, NELY, NELZ, l)
    K = fem_ref.assemble(NELX, NELY, NELZ, l, x)
    inter = lay.interior_dofs()  # i, k, j order == unpadded CSR order
    Kt = K[inter][:, inter].toarray()
    pos = {d: t for t, d in enumerate(inter)}
    fx = [pos[f] for f in lay.fixed_dofs()]
    Kt[fx, :] = 0
    Kt[:, fx] = 0
    Kt[fx, fx] = 1.0
    M = grid.csr(l, x).toarray()
    assert np.abs(M - Kt).max() <= 1e-13 * np.abs(Kt).max()


def test_filter_and_adjoint(grid):
    rmin = 1.5
    x = rand_density(9, halo=0.0)
    H = fem_ref.filter_matrix(NELX, NELY, NELZ, rmin)
    Hs = np.asarray(H.sum(axis=1)).ravel()
    lay0 = fem_ref.Layout(NELX, NELY, NELZ, 0)
    inter = np.array([fem_ref.fine_elem_index(lay0, i, j, k) for i in range(1, NELX + 1)
                      for k in range(1, NELZ + 1) for j in range(1, NELY + 1)])
    ref = np.zeros_like(x)
    ref[inter] = (H @ x)[inter] / Hs[inter]
    got = grid.filter(rmin, x)
    assert np.abs(got[inter] - ref[inter]).max() <= 1e-14
    v = rand_density(10, halo=0.0)
    g = grid.filter_grad(rmin, v)
    refg = H.T @ np.where(Hs > 0, v / np.where(Hs > 0, Hs, 1), 0)
    assert np.abs(g[inter] - refg[inter]).max() <= 1e-14
    # <F x, v> == <x, F^T v> on the interior
    assert np.dot(got[inter], v[inter]) == pytest.approx(np.dot(x[inter], g[inter]), rel=1e-13)
