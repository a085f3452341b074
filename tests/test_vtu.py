"""CPU checks of the density snapshot writers writeDensityVtkFile and
writeDensityVtkFileWithHalo (stencil_optimization.c:399-598).

The oracle's restatement is pinned by the structure of the file it writes:
XML, point coordinates, every cell the unit cube at its element corner, the
node order of each cell tied to the element matrix through its rigid-body
null space, offsets, types, and the densities read back in element order.
The library must then write the same bytes.  No GPU call is made."""
import ctypes
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import pyoracle
import top3d_amd

SHAPES = [(4, 4, 4), (8, 4, 12), (16, 8, 8)]


def padded_density(o, seed):
    # a distinct value on every padded element, so a transposed read fails
    return np.random.default_rng(seed).uniform(0.0, 1.0, o.nelem())


def parse(path):
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    arrays = {}
    for da in piece.iter("DataArray"):
        arrays[da.get("Name") or "points"] = np.array(da.text.split(), float)
    return root, piece, arrays


def rigid_modes(P):
    """6 rigid-body motions (3 translations, 3 rotations) of nodes at P (8x3)."""
    m = []
    for c in range(3):
        v = np.zeros((8, 3))
        v[:, c] = 1
        m.append(v.ravel())
    for a, b in ((0, 1), (1, 2), (2, 0)):
        v = np.zeros((8, 3))
        v[:, a], v[:, b] = -P[:, b], P[:, a]
        m.append(v.ravel())
    return np.array(m).T


def check_structure(path, x, o, halo):
    _, wx, wy, wz, _ = o.padding(0)
    px, py, pz = (wx, wy, wz) if halo else (o.nelx + 1, o.nely + 1, o.nelz + 1)
    ncell = (px - 1) * (py - 1) * (pz - 1)
    root, piece, arr = parse(path)
    assert root.tag == "VTKFile" and root.get("type") == "UnstructuredGrid"
    assert int(piece.get("NumberOfPoints")) == px * py * pz
    assert int(piece.get("NumberOfCells")) == ncell
    # points at the node indices, x slowest and y fastest
    pts = arr["points"].reshape(-1, 3)
    i, k, j = np.meshgrid(np.arange(px), np.arange(pz), np.arange(py), indexing="ij")
    assert np.array_equal(pts, np.stack([i.ravel(), j.ravel(), k.ravel()], 1))
    # cells in element order (x slowest, y fastest), each the unit cube at
    # its element corner
    conn = arr["connectivity"].astype(np.int64).reshape(ncell, 8)
    cp = pts[conn]
    lo = cp.min(axis=1)
    ci, ck, cj = np.meshgrid(np.arange(px - 1), np.arange(pz - 1), np.arange(py - 1),
                             indexing="ij")
    assert np.array_equal(lo, np.stack([ci.ravel(), cj.ravel(), ck.ravel()], 1))
    rel = np.unique(cp - lo[:, None, :], axis=0)
    assert rel.shape == (1, 8, 3)  # one local node order for all cells
    assert sorted(map(tuple, rel[0].astype(int))) == sorted(
        (a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1))
    # that order is the element matrix's: KE_0 annihilates the rigid motions
    # of the nodes at (x, -y, z) (KE_0 is integrated with y mirrored)
    KE = o.ke(0)[0]
    res = np.abs(KE @ rigid_modes(rel[0] * [1.0, -1.0, 1.0])).max()
    assert res <= 1e-12 * np.abs(KE).max()
    assert np.array_equal(arr["offsets"], 8 * np.arange(1, ncell + 1))
    assert np.all(arr["types"] == 12)
    # densities of the cells' elements; "%e" keeps 7 significant digits
    e0 = 0 if halo else 1
    ewy, ewz = wy - 1, wz - 1
    idx = ((ci + e0) * ewy * ewz + (ck + e0) * ewy + (cj + e0)).ravel()
    np.testing.assert_allclose(arr["density"], x[idx], rtol=6e-7, atol=0)


@pytest.mark.parametrize("halo", [False, True])
@pytest.mark.parametrize("shape", SHAPES)
def test_oracle_vtu_structure(tmp_path, shape, halo):
    o = pyoracle.Grid(*shape, 3)
    x = padded_density(o, 7)
    path = tmp_path / "oracle.vtu"
    o.write_vtu(x, str(path), halo)
    check_structure(str(path), x, o, halo)
    o.close()


def test_structure_check_catches_a_swapped_node_order(tmp_path):
    # the same file with local nodes 0 and 1 of every cell exchanged fails
    o = pyoracle.Grid(4, 4, 4, 3)
    x = padded_density(o, 3)
    path = tmp_path / "oracle.vtu"
    o.write_vtu(x, str(path), False)
    text = path.read_text().split("\n")
    a = text.index('<DataArray type="Int32" Name="connectivity" format= "ascii">') + 1
    for r in range(a, a + 64):
        q = text[r].split()
        q[0], q[1] = q[1], q[0]
        text[r] = " ".join(q)
    bad = tmp_path / "bad.vtu"
    bad.write_text("\n".join(text))
    with pytest.raises(AssertionError):
        check_structure(str(bad), x, o, False)
    o.close()


@pytest.mark.parametrize("halo", [False, True])
@pytest.mark.parametrize("shape", SHAPES)
def test_library_writes_the_oracles_bytes(tmp_path, shape, halo):
    o = pyoracle.Grid(*shape, 3)
    x = padded_density(o, 11)
    ref, got = tmp_path / "oracle.vtu", tmp_path / "lib.vtu"
    o.write_vtu(x, str(ref), halo)
    lib = top3d_amd.library()
    write = lib.writeDensityVtkFileWithHalo if halo else lib.writeDensityVtkFile
    write(*shape, x.ctypes.data_as(ctypes.c_void_p), str(got).encode())
    assert got.read_bytes() == ref.read_bytes()
    o.close()


def test_library_reports_an_unwritable_path(tmp_path):
    lib = top3d_amd.library()
    x = np.zeros(6 * 6 * 6)
    lib.top3d_clear_error()
    lib.writeDensityVtkFile(4, 4, 4, x.ctypes.data_as(ctypes.c_void_p),
                            str(tmp_path / "missing" / "x.vtu").encode())
    assert top3d_amd.ERRORS[lib.top3d_last_error()] == "argument"
    lib.top3d_clear_error()
