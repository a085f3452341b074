"""GPU parity of the level operators, transfers, diagonals, smoothers,
assembled coarse levels and design-loop kernels against the CPU oracle.
All calls go through the C ABI (libtop3d_hip.so via ctypes).  FP64 results
must agree to 1e-13 normwise (1e-12 where assembly order differs); FP32
kernels of the mixed V-cycle to FP32 rounding (1e-5)."""
import ctypes

import numpy as np
import pytest
import torch

import fem_ref
import pyoracle
import top3d_amd

pytestmark = pytest.mark.gpu

# (the last size spans several tiles in y and z at levels 0 and 1: 2 x 2 of
# the FP32 fine kernel's 33 x 17, 3 x 3 of the FP64 one's 17 x 12, 1 x 3 of
# level 1's 65 x 7 columns)
SIZES = [(16, 8, 8, 3), (24, 16, 8, 4), (40, 36, 28, 3)]


def relerr(got, ref):
    return np.abs(np.asarray(got) - np.asarray(ref)).max() / max(np.abs(ref).max(), 1e-300)


@pytest.fixture(scope="module", params=SIZES, ids=lambda s: "x".join(map(str, s)))
def pair(request):
    nelx, nely, nelz, nl = request.param
    print(f"[gpu ops] grid {nelx}x{nely}x{nelz} nl={nl}", flush=True)
    g = top3d_amd.Grid(nelx, nely, nelz, nl)
    o = pyoracle.Grid(nelx, nely, nelz, nl)
    assert g.padding(0) == o.padding(0)
    rng = np.random.default_rng(nelx)
    vals = rng.uniform(0.1, 1.0, nelx * nely * nelz)
    xg = g.density(vals)
    xo = xg.cpu().numpy()
    yield g, o, xg, xo, rng
    g.close()
    o.close()


def cuda(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0", dtype)


def test_fine_operator_fp64(pair):
    g, o, xg, xo, rng = pair
    v = rng.standard_normal(g.ndof(0))
    got = g.apply(xg, cuda(v)).cpu().numpy()
    ref = o.matvec0(xo, v)
    assert relerr(got, ref) <= 1e-13


def test_fine_operator_fp32(pair):
    g, o, xg, xo, rng = pair
    v = rng.standard_normal(g.ndof(0))
    got = g.apply_f32(xg, cuda(v, torch.float32)).cpu().numpy()
    ref = o.matvec0(xo, v.astype(np.float32).astype(np.float64))
    assert relerr(got, ref) <= 2e-6


def test_level1_operator(pair):
    g, o, xg, xo, rng = pair
    v = rng.standard_normal(g.ndof(1))
    got = g.apply_level(1, xg, cuda(v)).cpu().numpy()
    ref = o.matvec_l(1, xo, v)
    assert relerr(got, ref) <= 1e-13
    got32 = g.apply_level_f32(1, xg, cuda(v, torch.float32)).cpu().numpy()
    assert relerr(got32, o.matvec_l(1, xo, v.astype(np.float32).astype(np.float64))) <= 2e-6


def test_level2_operator_identity_rows(pair):
    g, o, xg, xo, rng = pair
    v = rng.standard_normal(g.ndof(2))
    got = g.apply_level(2, xg, cuda(v)).cpu().numpy()
    ref = o.matvec_l(2, xo, v)
    assert relerr(got, ref) <= 1e-12


def test_transfers(pair):
    g, o, xg, xo, rng = pair
    for l in range(g.nl - 1):
        vc = rng.standard_normal(g.ndof(l + 1))
        assert relerr(g.prolong(l, cuda(vc)).cpu().numpy(), o.prolong(l, vc)) <= 1e-15
        vf = rng.standard_normal(g.ndof(l))
        assert relerr(g.restrict(l, cuda(vf)).cpu().numpy(), o.restrict(l, vf)) <= 1e-15


def test_inverse_diagonals(pair):
    g, o, xg, xo, rng = pair
    for l in range(g.nl):
        assert relerr(g.invdiag(xg, l).cpu().numpy(), o.invdiag(xo, l)) <= 1e-13


def test_jacobi_smoothers(pair):
    g, o, xg, xo, rng = pair
    for l in (0, 1):
        inv = o.invdiag(xo, l)
        u = rng.standard_normal(g.ndof(l))
        b = rng.standard_normal(g.ndof(l))
        ug = g.jacobi(xg, 3, cuda(inv), cuda(u), cuda(b), l=l).cpu().numpy()
        uo = u.copy()
        tmp = np.zeros_like(u)
        P = ctypes.c_void_p
        ptrs = [P(inv.ctypes.data), P(uo.ctypes.data), P(b.ctypes.data), P(tmp.ctypes.data)]
        if l == 0:
            fn = o.lib.or_jacobi0
            fn.argtypes = [P, P, ctypes.c_int, ctypes.c_double, P, P, P, P]
            fn(P(o.h), P(xo.ctypes.data), 3, 0.6, *ptrs)
        else:
            fn = o.lib.or_jacobi_l
            fn.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_double, P, P, P, P]
            fn(P(o.h), P(xo.ctypes.data), l, 3, 0.6, *ptrs)
        assert relerr(ug, uo) <= 1e-12


def test_assembled_level_csr(pair):
    g, o, xg, xo, rng = pair
    lib = g.lib
    for l in range(2, g.nl):
        M = top3d_amd.CSRMatrix()
        lib.allocateSubspaceMatrix(g.gc, l, ctypes.byref(M))
        tmp = np.zeros(M.nrows)
        lib.assembleSubspaceMatrix(g.gc, l, top3d_amd._ptr(xg), M,
                                   tmp.ctypes.data_as(ctypes.c_void_p))
        top3d_amd.check("assembleSubspaceMatrix")
        vals = np.ctypeslib.as_array(M.vals, (M.nnz,)).copy()
        ref = o.csr(l, xo)
        assert relerr(vals, ref.data) <= 1e-12
        # CSR SpMV in the halo layout
        v = np.zeros(g.ndof(l))
        lay = fem_ref.Layout(g.nelx, g.nely, g.nelz, l)
        inter = lay.interior_dofs()
        v[inter] = rng.standard_normal(inter.size)
        out = g.zeros(l)
        lib.applyStateOperatorSubspaceMatrix(g.gc, l, M, top3d_amd._ptr(cuda(v)),
                                             top3d_amd._ptr(out))
        top3d_amd.check("applyStateOperatorSubspaceMatrix")
        refo = np.zeros(g.ndof(l))
        refo[inter] = ref @ v[inter]
        assert relerr(out.cpu().numpy(), refo) <= 1e-12
        lib.freeSubspaceMatrix(ctypes.byref(M))


def test_coarse_direct_solve(pair):
    g, o, xg, xo, rng = pair
    lib = g.lib
    l = g.nl - 1
    M = top3d_amd.CSRMatrix()
    lib.allocateSubspaceMatrix(g.gc, l, ctypes.byref(M))
    lib.assembleSubspaceMatrix(g.gc, l, top3d_amd._ptr(xg), M, None)
    cs = top3d_amd.CoarseSolverData()
    lib.initializeCholmod(g.gc, l, ctypes.byref(cs), M)
    lib.factorizeSubspaceMatrix(g.gc, l, cs, M)
    top3d_amd.check("factorizeSubspaceMatrix")
    lay = fem_ref.Layout(g.nelx, g.nely, g.nelz, l)
    rhs = np.zeros(g.ndof(l))
    rhs[lay.interior_dofs()] = rng.standard_normal(lay.interior_dofs().size)
    out = g.zeros(l)
    lib.solveSubspaceMatrix(g.gc, l, cs, top3d_amd._ptr(cuda(rhs)), top3d_amd._ptr(out))
    top3d_amd.check("solveSubspaceMatrix")
    assert o.setup(xo) == 0
    ref = o.coarse_solve(rhs)
    assert relerr(out.cpu().numpy(), ref) <= 1e-10
    lib.finishCholmod(g.gc, l, ctypes.byref(cs), M, 0)
    lib.freeSubspaceMatrix(ctypes.byref(M))


def test_sensitivity_filter_volume_change_oc(pair):
    g, o, xg, xo, rng = pair
    u = rng.standard_normal(g.ndof(0))
    c, dc = g.sensitivity(xg, cuda(u))
    co, dco = o.sensitivity(xo, u)
    inter = g.interior_elements()
    assert abs(c - co) <= 1e-12 * abs(co)
    assert relerr(dc.cpu().numpy()[inter], dco[inter]) <= 1e-12
    rmin = 1.5
    assert relerr(g.filter(rmin, xg).cpu().numpy()[inter], o.filter(rmin, xo)[inter]) <= 1e-14
    v = g.density(rng.uniform(-1, 1, inter.size))
    assert relerr(g.filter_grad(rmin, v).cpu().numpy()[inter],
                  o.filter_grad(rmin, v.cpu().numpy())[inter]) <= 1e-14
    assert g.volume(xg) == pytest.approx(o.volume(xo), rel=1e-13)
    dv = g.filter_grad(rmin, torch.ones_like(xg))
    dcg = g.filter_grad(rmin, dc)
    xnew = torch.zeros_like(xg)
    g.oc(xg, xnew, dv, dcg, 0.01)
    ref_new = o.oc(xo, dv.cpu().numpy(), dcg.cpu().numpy(), 0.01)
    assert np.abs(xnew.cpu().numpy() - ref_new).max() <= 1e-12
    xc = xg.clone()
    ch = g.change(xc, xnew)
    assert ch == pytest.approx(float(np.float32(np.abs(xo - ref_new).max())), rel=1e-6)
    assert torch.equal(xc, xnew)


@pytest.mark.parametrize("rmin", [1.2, 2.0, 2.2, 2.5, 4.5])
def test_filter_radii(pair, rmin):
    """Filter and its transpose at radii with 7 and 27 non-zero window weights
    (the fixed-window kernel for radii up to 2, as 1.5 in the other tests), 33
    and 81 (the tap-list kernels of 64 and 240 taps) and more than 240 (the
    full-window kernel)."""
    g, o, xg, xo, rng = pair
    inter = g.interior_elements()
    assert relerr(g.filter(rmin, xg).cpu().numpy()[inter], o.filter(rmin, xo)[inter]) <= 1e-14
    v = g.density(rng.uniform(-1, 1, inter.size))
    assert relerr(g.filter_grad(rmin, v).cpu().numpy()[inter],
                  o.filter_grad(rmin, v.cpu().numpy())[inter]) <= 1e-14


def test_reference_layout_adapter_operator_and_solve():
    """Arrays in the reference's padded layout (y to a multiple of 128,
    gpu_definitions.c:67-92) go through the adapter into the library and
    back: the fine operator matches the oracle built with that padding to
    1e-13, and a solve on its load and densities gives its displacement."""
    nelx, nely, nelz, nl = 16, 8, 8, 3
    g = top3d_amd.Grid(nelx, nely, nelz, nl)
    o = pyoracle.Grid(nelx, nely, nelz, nl, yalign=128)
    lib, gc = g.lib, ctypes.byref(g.gc)
    P = ctypes.c_void_p
    rng = np.random.default_rng(9)
    _, wx, wy, wz, ndof_ref = o.padding(0)
    i, k, j = np.meshgrid(np.arange(1, nelx + 1), np.arange(1, nelz + 1),
                          np.arange(1, nely + 1), indexing="ij")
    x_ref = np.zeros(o.nelem())
    x_ref[(i * (wy - 1) * (wz - 1) + k * (wy - 1) + j).ravel()] = rng.uniform(
        0.1, 1.0, nelx * nely * nelz)
    v_ref = np.zeros(ndof_ref)
    node = np.zeros((wx, wz, wy), bool)
    node[1:nelx + 2, 1:nelz + 2, 1:nely + 2] = True
    v_ref[np.repeat(node.ravel(), 3)] = rng.standard_normal(3 * node.sum())
    x_lib = np.zeros(g.nelem)
    v_lib = np.zeros(g.ndof(0))
    lib.top3d_elements_from_reference_layout(gc, P(x_ref.ctypes.data), P(x_lib.ctypes.data))
    lib.top3d_nodes_from_reference_layout(gc, 0, P(v_ref.ctypes.data), P(v_lib.ctypes.data))
    out_lib = g.apply(cuda(x_lib), cuda(v_lib)).cpu().numpy()
    out_ref = np.zeros(ndof_ref)
    lib.top3d_nodes_to_reference_layout(gc, 0, P(out_lib.ctypes.data), P(out_ref.ctypes.data))
    want = o.matvec0(x_ref, v_ref)
    real = np.repeat(node.ravel(), 3)
    assert relerr(out_ref[real], want[real]) <= 1e-13
    # the solve through the adapter: the oracle's load and uniform densities
    F_ref = o.load()
    xu_ref = np.where(x_ref > 0, float(np.float32(0.2)), 0.0)
    F_lib, xu_lib = np.zeros(g.ndof(0)), np.zeros(g.nelem)
    lib.top3d_nodes_from_reference_layout(gc, 0, P(F_ref.ctypes.data), P(F_lib.ctypes.data))
    lib.top3d_elements_from_reference_layout(gc, P(xu_ref.ctypes.data), P(xu_lib.ctypes.data))
    U = g.zeros(0)
    it, _ = g.solve(cuda(xu_lib), cuda(F_lib), U)
    U_ref = np.zeros(ndof_ref)
    U_host = U.cpu().numpy()
    lib.top3d_nodes_to_reference_layout(gc, 0, P(U_host.ctypes.data), P(U_ref.ctypes.data))
    Uo = np.zeros(ndof_ref)
    ito, _ = o.solve(xu_ref, F_ref, Uo)
    assert it == ito
    assert np.linalg.norm(U_ref - Uo) <= 1e-9 * np.linalg.norm(Uo)
    g.close()
    o.close()


def test_coarse_factor_at_production_size():
    """The coarsest level of every benchmark configuration (coarse 16x8x8,
    n = 4131 DOFs, nested-dissection tiled Cholesky and W = L^-1 as the
    solver uses them) through initializeCholmod / factorizeSubspaceMatrix /
    solveSubspaceMatrix (stencil_assembly.c:366-503), against a numpy dense
    solve of the same CSR matrix, on random densities in [0.1, 1]."""
    nelx, nely, nelz, nl = 64, 32, 32, 3
    print("[gpu coarse] 64x32x32 nl=3: coarse n = 4131", flush=True)
    g = top3d_amd.Grid(nelx, nely, nelz, nl)
    lib = g.lib
    rng = np.random.default_rng(21)
    xg = g.density(rng.uniform(0.1, 1.0, nelx * nely * nelz))
    l = nl - 1
    M = top3d_amd.CSRMatrix()
    lib.allocateSubspaceMatrix(g.gc, l, ctypes.byref(M))
    lib.assembleSubspaceMatrix(g.gc, l, top3d_amd._ptr(xg), M, None)
    top3d_amd.check("assembleSubspaceMatrix")
    n = M.nrows
    assert n == 4131
    ro = np.ctypeslib.as_array(M.rowOffsets, (n + 1,)).copy()
    ci = np.ctypeslib.as_array(M.colIndex, (M.nnz,)).copy()
    va = np.ctypeslib.as_array(M.vals, (M.nnz,)).copy()
    A = np.zeros((n, n))
    for r in range(n):
        A[r, ci[ro[r]:ro[r + 1]]] = va[ro[r]:ro[r + 1]]
    cs = top3d_amd.CoarseSolverData()
    lib.initializeCholmod(g.gc, l, ctypes.byref(cs), M)
    lib.factorizeSubspaceMatrix(g.gc, l, cs, M)
    top3d_amd.check("factorizeSubspaceMatrix")
    lay = fem_ref.Layout(nelx, nely, nelz, l)
    inter = lay.interior_dofs()
    assert inter.size == n
    for trial in range(2):
        rhs = np.zeros(g.ndof(l))
        rhs[inter] = rng.standard_normal(n)
        out = g.zeros(l)
        lib.solveSubspaceMatrix(g.gc, l, cs, top3d_amd._ptr(cuda(rhs)), top3d_amd._ptr(out))
        top3d_amd.check("solveSubspaceMatrix")
        want = np.linalg.solve(A, rhs[inter])
        got = out.cpu().numpy()[inter]
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(f"[gpu coarse] trial {trial}: rel err {err:.2e}", flush=True)
        assert err <= 1e-10
    lib.finishCholmod(g.gc, l, ctypes.byref(cs), M, 0)
    lib.freeSubspaceMatrix(ctypes.byref(M))
    g.close()
