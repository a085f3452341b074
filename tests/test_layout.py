"""The reference-layout adapter (top3d_reference_padding,
top3d_{nodes,elements}_{from,to}_reference_layout) on the CPU: the
reference's padding rule (gpu_definitions.c:67-92, y padded to a multiple of
STENCIL_SIZE_Y = 128) against the oracle built with that alignment, the
cantilever load and seeded element fields carried between the two layouts,
and round trips.  No GPU call is made."""
import ctypes

import numpy as np
import pytest

import pyoracle
import top3d_amd

P = ctypes.c_void_p


def grid_ctx(nelx, nely, nelz, nl):
    """The fields the layout functions read (setupGC would also build the
    device grid)."""
    lib = top3d_amd.library()
    gc = top3d_amd.GridContext()
    gc.nelx, gc.nely, gc.nelz = nelx, nely, nelz
    return lib, gc


def ref_padding(lib, gc, l):
    v = [ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_uint64()]
    lib.top3d_reference_padding(ctypes.byref(gc), l, *[ctypes.byref(q) for q in v])
    return tuple(q.value for q in v)


@pytest.mark.parametrize("shape", [(16, 8, 8, 3), (24, 16, 8, 4), (32, 256, 16, 3)])
def test_reference_padding_matches_oracle(shape):
    nelx, nely, nelz, nl = shape
    lib, gc = grid_ctx(*shape)
    o = pyoracle.Grid(nelx, nely, nelz, nl, yalign=128)
    for l in range(nl):
        _, wx, wy, wz, ndof = o.padding(l)
        assert ref_padding(lib, gc, l) == (wx, wy, wz, ndof)
        assert (wy - 3 + 1) % 128 == 0 or (nely >> l) + 1 == 128
    o.close()


@pytest.mark.parametrize("shape", [(16, 8, 8, 3), (24, 16, 8, 4)])
def test_load_and_densities_carried_between_layouts(shape):
    nelx, nely, nelz, nl = shape
    lib, gc = grid_ctx(*shape)
    oref = pyoracle.Grid(nelx, nely, nelz, nl, yalign=128)
    olib = pyoracle.Grid(nelx, nely, nelz, nl, yalign=1)
    Fref, Flib = oref.load(), olib.load()
    got = np.full(Flib.size, np.nan)
    lib.top3d_nodes_from_reference_layout(ctypes.byref(gc), 0, P(Fref.ctypes.data),
                                          P(got.ctypes.data))
    top3d_amd.check("nodes_from_reference_layout")
    assert np.array_equal(got, Flib)
    back = np.full(Fref.size, np.nan)
    lib.top3d_nodes_to_reference_layout(ctypes.byref(gc), 0, P(got.ctypes.data),
                                        P(back.ctypes.data))
    assert np.array_equal(back, Fref)
    # a distinct value on every real element, the halo and padding zero
    rng = np.random.default_rng(3)
    vals = rng.uniform(0.1, 1.0, nelx * nely * nelz)
    xr = np.zeros(oref.nelem())
    _, _, wyr, wzr, _ = oref.padding(0)
    _, _, wyl, wzl, _ = olib.padding(0)
    i, k, j = np.meshgrid(np.arange(1, nelx + 1), np.arange(1, nelz + 1),
                          np.arange(1, nely + 1), indexing="ij")
    xr[(i * (wyr - 1) * (wzr - 1) + k * (wyr - 1) + j).ravel()] = vals
    xl = np.full(olib.nelem(), np.nan)
    lib.top3d_elements_from_reference_layout(ctypes.byref(gc), P(xr.ctypes.data),
                                             P(xl.ctypes.data))
    want = np.zeros(olib.nelem())
    want[(i * (wyl - 1) * (wzl - 1) + k * (wyl - 1) + j).ravel()] = vals
    assert np.array_equal(xl, want)
    xr2 = np.full(xr.size, np.nan)
    lib.top3d_elements_to_reference_layout(ctypes.byref(gc), P(xl.ctypes.data),
                                           P(xr2.ctypes.data))
    assert np.array_equal(xr2, xr)
    # level vectors: a round trip keeps every entry of the library layout
    for l in range(1, nl):
        n = olib.ndof(l)
        v = rng.standard_normal(n)
        r = np.full(ref_padding(lib, gc, l)[3], np.nan)
        lib.top3d_nodes_to_reference_layout(ctypes.byref(gc), l, P(v.ctypes.data),
                                            P(r.ctypes.data))
        assert not np.isnan(r).any()
        w = np.full(n, np.nan)
        lib.top3d_nodes_from_reference_layout(ctypes.byref(gc), l, P(r.ctypes.data),
                                              P(w.ctypes.data))
        assert np.array_equal(w, v)
    oref.close()
    olib.close()
