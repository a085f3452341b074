"""Supports and loads beyond the shipped cantilever, on the CPU.

The reference builds two clamped strips of the x=0 face (setFixedDof_halo,
stencil_grid_utility.c:34-57) and keeps the whole-face clamp of its MATLAB
specification in comments (:19-32; top3dmgcg_reference.m:62).  The MATLAB
line load (top3dmgcg_reference.m:59) puts F_z = -1 on every node of the edge
x = nelx, z = 0.  The oracle's versions are pinned by their node sets and
totals; the library's host-side lists must equal them entry by entry.  No GPU
call is made."""
import ctypes

import numpy as np
import pytest

import pyoracle
import top3d_amd

SHAPE = (16, 8, 12, 3)


def node_coords(o, l, dofs):
    _, wx, wy, wz, _ = o.padding(l)
    n = dofs // 3
    return n // (wy * wz), (n % (wy * wz)) % wy, (n % (wy * wz)) // wy, dofs % 3  # i, j, k, c


@pytest.mark.parametrize("l", [0, 1, 2])
def test_oracle_face_support_is_the_whole_face(l):
    o = pyoracle.Grid(*SHAPE, support="face")
    s = pyoracle.Grid(*SHAPE)
    ny, nz = SHAPE[1] // 2 ** l + 1, SHAPE[2] // 2 ** l + 1
    f = o.fixed(l)
    assert f.size == 3 * ny * nz and np.unique(f).size == f.size
    i, j, k, c = node_coords(o, l, f)
    assert np.all(i == 1)
    got = set(zip(j.tolist(), k.tolist(), c.tolist()))
    assert got == {(a, b, q) for a in range(1, ny + 1) for b in range(1, nz + 1) for q in range(3)}
    # the shipped strips are part of the face
    assert set(s.fixed(l).tolist()) < set(f.tolist())
    o.close()
    s.close()


def test_oracle_line_load():
    o = pyoracle.Grid(*SHAPE)
    F = o.line_load()
    nz = np.flatnonzero(F)
    assert nz.size == SHAPE[1] + 1 and F.sum() == -(SHAPE[1] + 1)
    i, j, k, c = node_coords(o, 0, nz)
    assert np.all(i == SHAPE[0] + 1) and np.all(k == 1) and np.all(c == 2)
    assert sorted(j.tolist()) == list(range(1, SHAPE[1] + 2))
    o.close()


def host_context(nelx, nely, nelz, nl):
    """A gridContext with sizes, wraps and an empty fixed-DOF table only."""
    lib = top3d_amd.library()
    gc = top3d_amd.GridContext()
    gc.nelx, gc.nely, gc.nelz = nelx, nely, nelz
    v = [ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_uint64()]
    lib.computePadding(ctypes.byref(gc), 0, *[ctypes.byref(q) for q in v])
    gc.wrapx, gc.wrapy, gc.wrapz = v[1].value, v[2].value, v[3].value
    table = (top3d_amd.FixedDofs * nl)()
    gc.fixedDofs = ctypes.cast(table, ctypes.POINTER(top3d_amd.FixedDofs))
    return gc, table


@pytest.mark.parametrize("support", ["strips", "face"])
def test_library_fixed_dofs_match_oracle(support):
    lib = top3d_amd.library()
    gc, table = host_context(*SHAPE)
    o = pyoracle.Grid(*SHAPE, support=support)
    prev = lib.top3d_get_default_support()
    lib.top3d_set_default_support(top3d_amd.SUPPORTS[support])
    try:
        for l in range(SHAPE[3]):
            lib.setFixedDof_halo(ctypes.byref(gc), l)
            fd = table[l]
            got = np.ctypeslib.as_array(fd.idx, (fd.n,)).astype(np.int64)
            assert np.array_equal(got, o.fixed(l))
    finally:
        lib.top3d_set_default_support(prev)
    o.close()


def test_library_line_load_matches_oracle():
    lib = top3d_amd.library()
    gc, _ = host_context(*SHAPE)
    o = pyoracle.Grid(*SHAPE)
    F = np.zeros(o.ndof(0))
    lib.top3d_line_load(ctypes.byref(gc), F.ctypes.data_as(ctypes.c_void_p))
    assert np.array_equal(F, o.line_load())
    o.close()


@pytest.mark.parametrize("profile", [1, 2])
def test_oracle_line_load_profiles(profile):
    # end-weighted: total nely; sine: sum_j sin(pi j / n) = cot(pi / (2 n)),
    # zero at both ends and mirror-symmetric about the middle node
    o = pyoracle.Grid(*SHAPE)
    F, F0 = o.line_load(profile), o.line_load(0)
    n = SHAPE[1]
    nodes = np.flatnonzero(F0)  # the edge's z components, j = 0..nely in order
    w = -F[nodes]
    assert np.all(F[np.setdiff1d(np.arange(F.size), nodes)] == 0)
    if profile == 1:
        assert w.sum() == n and w[0] == w[-1] == 0.5 and np.all(w[1:-1] == 1)
    else:
        assert abs(w.sum() - 1 / np.tan(np.pi / (2 * n))) <= 1e-13 * n
        assert abs(w[0]) <= 1e-15 and abs(w[-1]) <= 1e-15
        np.testing.assert_allclose(w, w[::-1], rtol=0, atol=1e-15)
        assert w[n // 2] == pytest.approx(1.0, abs=1e-15)
    o.close()


@pytest.mark.parametrize("profile", [0, 1, 2])
def test_library_line_load_profiles_match_oracle(profile):
    lib = top3d_amd.library()
    gc, _ = host_context(*SHAPE)
    o = pyoracle.Grid(*SHAPE)
    F = np.zeros(o.ndof(0))
    lib.top3d_line_load_profile(ctypes.byref(gc), F.ctypes.data_as(ctypes.c_void_p), profile)
    np.testing.assert_allclose(F, o.line_load(profile), rtol=1e-15, atol=1e-17)
    o.close()
