"""CPU checks of the C ABI: the library loads and exports every entry point
declared in include/top3d_hip.h, the structs keep the reference's layout, and
the host-only entry points (getKEsubspace, computePadding,
allocateSubspaceMatrix) agree with the oracle.  No GPU call is made."""
import ctypes
import os
import re

import numpy as np
import pytest

import pyoracle
import top3d_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "top3d_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("defined",)))


def test_library_exports_every_declared_symbol():
    lib = top3d_amd.library()
    names = declared_functions()
    assert len(names) > 50
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the binding declares a signature for each of them
    assert set(names) <= set(top3d_amd.SIGNATURES)


def test_struct_layout_matches_reference():
    # uint_fast32_t is 8 bytes on x86-64 glibc (definitions.h:31-53)
    assert ctypes.sizeof(top3d_amd.FixedDofs) == 16
    assert top3d_amd.GridContext.nelx.offset == 48
    assert top3d_amd.GridContext.penal.offset == 96
    assert ctypes.sizeof(top3d_amd.GridContext) == 120
    assert ctypes.sizeof(top3d_amd.CSRMatrix) == 40
    assert ctypes.sizeof(top3d_amd.CoarseSolverData) == 56
    assert ctypes.sizeof(top3d_amd.SolverDataBuffer) == 8 * 9 + 56
    assert ctypes.sizeof(top3d_amd.GpuNode) == 40
    assert ctypes.sizeof(top3d_amd.GpuGrid) == 16


@pytest.mark.parametrize("l", [0, 1, 2, 3])
def test_library_ke_matches_oracle(l):
    lib = top3d_amd.library()
    n = 1 << l
    got = np.zeros(576 * n ** 3)
    lib.getKEsubspace(got.ctypes.data_as(ctypes.c_void_p), 0.3, l)
    ref = np.zeros(576 * n ** 3)
    pyoracle.load().or_ke(0.3, l, 10, ref.ctypes.data_as(ctypes.c_void_p))
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


def test_compute_padding_matches_oracle_layout():
    lib = top3d_amd.library()
    gc = top3d_amd.GridContext()
    gc.nelx, gc.nely, gc.nelz = 64, 32, 32
    o = pyoracle.Grid(64, 32, 32, 3)
    for l in range(3):
        v = [ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_uint64()]
        lib.computePadding(ctypes.byref(gc), l, *[ctypes.byref(q) for q in v])
        assert tuple(q.value for q in v) == o.padding(l)
    o.close()


def test_csr_pattern_matches_oracle():
    lib = top3d_amd.library()
    gc = top3d_amd.GridContext()
    gc.nelx, gc.nely, gc.nelz = 16, 8, 8
    M = top3d_amd.CSRMatrix()
    lib.allocateSubspaceMatrix(gc, 2, ctypes.byref(M))
    o = pyoracle.Grid(16, 8, 8, 3)
    ref = o.csr(2, np.zeros(o.nelem()))
    n, nnz = M.nrows, M.nnz
    ro = np.ctypeslib.as_array(M.rowOffsets, (n + 1,)).copy()
    ci = np.ctypeslib.as_array(M.colIndex, (nnz,)).copy()
    assert np.array_equal(ro, ref.indptr) and np.array_equal(ci, ref.indices)
    lib.freeSubspaceMatrix(ctypes.byref(M))
    o.close()


# Layout of the reference's structs on x86-64 glibc, where uint_fast32_t is
# 8 bytes: definitions.h:31-53, gpu_definitions.h:7-37 and
# stencil_assembly.h:8-27 (CHOLMOD handles are pointers).
REFERENCE_LAYOUT = {
    "FixedDofs": 16, "FixedDofs.n": 0, "FixedDofs.idx": 8,
    "gridContext": 120, "gridContext.E0": 0, "gridContext.Emin": 8, "gridContext.nu": 16,
    "gridContext.elementSizeX": 24, "gridContext.elementSizeY": 32,
    "gridContext.elementSizeZ": 40, "gridContext.nelx": 48, "gridContext.nely": 56,
    "gridContext.nelz": 64, "gridContext.wrapx": 72, "gridContext.wrapy": 80,
    "gridContext.wrapz": 88, "gridContext.penal": 96, "gridContext.precomputedKE": 104,
    "gridContext.fixedDofs": 112,
    "CSRMatrix": 40, "CSRMatrix.nnz": 0, "CSRMatrix.nrows": 8, "CSRMatrix.rowOffsets": 16,
    "CSRMatrix.colIndex": 24, "CSRMatrix.vals": 32,
    "CoarseSolverData": 56, "CoarseSolverData.cholmodCommon": 0,
    "CoarseSolverData.factoredMatrix": 16, "CoarseSolverData.E_workspace": 48,
    "SolverDataBuffer": 128, "SolverDataBuffer.r": 0, "SolverDataBuffer.p": 8,
    "SolverDataBuffer.q": 16, "SolverDataBuffer.z": 24, "SolverDataBuffer.invD": 32,
    "SolverDataBuffer.dmg": 40, "SolverDataBuffer.rmg": 48, "SolverDataBuffer.zmg": 56,
    "SolverDataBuffer.coarseMatrices": 64, "SolverDataBuffer.bottomSolver": 72,
    "gpuNode": 40, "gpuNode.id": 0, "gpuNode.prev": 8, "gpuNode.next": 16, "gpuNode.gc": 24,
    "gpuNode.offset_x": 32, "gpuNode.x_global": 36,
    "gpuGrid": 16, "gpuGrid.num_targets": 0, "gpuGrid.target": 8,
    "uint_fast32_t": 8,
}


def test_header_compiled_as_c_has_reference_layout(tmp_path):
    """include/top3d_hip.h compiled by a C compiler (as the reference's
    callers are) lays out every struct as the reference does, and the
    Python binding's ctypes mirror agrees."""
    import subprocess
    exe = tmp_path / "abi_layout"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "native", "abi_layout.c"), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {k: int(v) for k, v in (line.split() for line in out.splitlines())}
    assert got == REFERENCE_LAYOUT
    mirror = {"FixedDofs": top3d_amd.FixedDofs, "gridContext": top3d_amd.GridContext,
              "CSRMatrix": top3d_amd.CSRMatrix, "CoarseSolverData": top3d_amd.CoarseSolverData,
              "SolverDataBuffer": top3d_amd.SolverDataBuffer, "gpuNode": top3d_amd.GpuNode,
              "gpuGrid": top3d_amd.GpuGrid}
    for key, val in got.items():
        if key == "uint_fast32_t":
            continue
        s, _, field = key.partition(".")
        cls = mirror[s]
        assert (getattr(cls, field).offset if field else ctypes.sizeof(cls)) == val, key
