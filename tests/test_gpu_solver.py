"""GPU parity of the MGCG state solve and the design loop against the CPU
oracle, through the C ABI.

Tolerances: FP64 form -- same CG iteration count and
per-iteration compliance within 1e-8 relative of the FP64 oracle; mixed form
(FP32 V-cycle, FP64 outer CG) -- compliance within 1e-5 relative."""
import json
import os

import numpy as np
import pytest
import torch

import fem_ref
import pyoracle
import top3d_amd

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def uniform(g):
    return g.density(fill=float(np.float32(0.2)))


@pytest.mark.parametrize("shape", [(16, 8, 8, 3), (32, 16, 16, 4), (24, 8, 16, 3)])
def test_fp64_solve_matches_oracle(shape):
    nelx, nely, nelz, nl = shape
    print(f"[gpu solve] {shape}", flush=True)
    g = top3d_amd.Grid(nelx, nely, nelz, nl)
    o = pyoracle.Grid(nelx, nely, nelz, nl)
    F = g.load()
    x = uniform(g)
    U = g.zeros(0)
    it, res = g.solve(x, F, U)
    Uo = np.zeros(o.ndof(0))
    ito, reso = o.solve(x.cpu().numpy(), F.cpu().numpy(), Uo)
    assert it == ito
    assert res == pytest.approx(reso, rel=1e-6)
    Ug = U.cpu().numpy()
    assert np.linalg.norm(Ug - Uo) <= 1e-9 * np.linalg.norm(Uo)
    c, co = F.cpu().numpy() @ Ug, F.cpu().numpy() @ Uo
    assert abs(c - co) <= 1e-10 * abs(co)
    # warm start: the next solve from the converged U stops at iteration 0
    it2, _ = g.solve(x, F, U)
    assert it2 == 0
    g.close()
    o.close()


def test_symmetric_bc_random_density_converges():
    nelx, nely, nelz, nl = 16, 8, 8, 3
    g = top3d_amd.Grid(nelx, nely, nelz, nl, bc="symmetric")
    o = pyoracle.Grid(nelx, nely, nelz, nl, bcsym=True)
    rng = np.random.default_rng(5)
    x = g.density(rng.uniform(0.1, 1.0, nelx * nely * nelz))
    F = g.load()
    U = g.zeros(0)
    it, res = g.solve(x, F, U, tol=1e-10)
    Uo = np.zeros(o.ndof(0))
    ito, reso = o.solve(x.cpu().numpy(), F.cpu().numpy(), Uo, tol=1e-10)
    # ~150 iterations to 1e-10: rounding-order differences move the count a
    # little, the converged solutions agree
    assert res < 1e-10 and abs(it - ito) <= max(2, ito // 10)
    assert np.linalg.norm(U.cpu().numpy() - Uo) <= 1e-9 * np.linalg.norm(Uo)
    g.close()
    o.close()


def test_reference_bc_stagnation_is_reported():
    # identity-row BCs stagnate on heterogeneous densities (as the oracle
    # does); the solver reports TOP3D_ERR_NO_CONVERGE instead of hanging
    g = top3d_amd.Grid(16, 8, 8, 3)
    g.set_max_iter(50)
    rng = np.random.default_rng(5)
    x = g.density(rng.uniform(0.1, 1.0, 16 * 8 * 8))
    U = g.zeros(0)
    with pytest.raises(top3d_amd.Top3dError, match="no convergence"):
        g.solve(x, g.load(), U)
    g.close()


def test_mixed_solve_accuracy():
    nelx, nely, nelz, nl = 32, 16, 16, 4
    g64 = top3d_amd.Grid(nelx, nely, nelz, nl)
    g32 = top3d_amd.Grid(nelx, nely, nelz, nl, precision="mixed")
    F = g64.load()
    x = uniform(g64)
    U64, U32 = g64.zeros(0), g32.zeros(0)
    it64, r64 = g64.solve(x, F, U64)
    it32, r32 = g32.solve(x, F, U32)
    assert r32 < 1e-5 and abs(it32 - it64) <= 2
    c64 = float(F @ U64)
    c32 = float(F @ U32)
    assert abs(c32 - c64) <= 1e-5 * abs(c64)
    g64.close()
    g32.close()


@pytest.mark.parametrize("precision,typing", [("fp64", "fp64"), ("mixed", "fp64")])
def test_design_loop_c1_matches_oracle_trajectory(precision, typing):
    ref = json.load(open(os.path.join(GOLDEN, f"c1_{typing}.json")))
    g = top3d_amd.Grid(ref["nelx"], ref["nely"], ref["nelz"], ref["nl"], precision=precision)
    recs, st = top3d_amd.design_loop(g, ref["iters"])
    assert len(recs) == len(ref["records"])
    tol = 1e-8 if precision == "fp64" else 1e-5
    for a, b in zip(recs, ref["records"]):
        assert abs(a["c"] - b["c"]) <= tol * abs(b["c"]), (a, b)
        assert abs(a["vol"] - b["vol"]) <= 1e-6
        if precision == "fp64":
            assert a["cgiter"] == b["cgiter"]
    if precision == "fp64":
        xp = np.load(os.path.join(GOLDEN, "c1_xphys_fp64.npz"))["xphys"]
        got = st.xPhys.cpu().numpy()[g.interior_elements()]
        assert np.abs(got - xp).max() <= 1e-6
    g.close()


def test_autograd_compliance_gradient_fd():
    # 16x8x8 stand-in for the 64x32x32 differentiation case: FP64,
    # symmetric BCs, tight tolerance, gradient through the filter vs central
    # differences
    nelx, nely, nelz, nl = 16, 8, 8, 3
    g = top3d_amd.Grid(nelx, nely, nelz, nl, bc="symmetric")
    g.set_max_iter(2000)
    vals = fem_ref.splitmix64_densities(nelx * nely * nelz)
    x = g.density(vals)
    F = g.load()
    rmin = 1.5

    def compliance(xx):
        xp = g.filter(rmin, xx)
        return top3d_amd.Compliance.apply(xp, g, F, tol=1e-12)

    xr = x.clone().requires_grad_(True)
    xphys = g.filter(rmin, xr.detach())
    c = top3d_amd.Compliance.apply(xphys.requires_grad_(True), g, F, 1e-12, 0.0)
    c.backward()
    grad_xphys = xphys.grad
    # chain through the filter by the library's transposed filter
    grad_x = g.filter_grad(rmin, grad_xphys)
    idx = g.interior_elements()
    rng = np.random.default_rng(3)
    pick = rng.choice(idx, 8, replace=False)
    h = 1e-4
    for e in pick:
        xp_, xm_ = x.clone(), x.clone()
        xp_[e] += h
        xm_[e] -= h
        fd = (float(compliance(xp_)) - float(compliance(xm_))) / (2 * h)
        assert float(grad_x[e]) == pytest.approx(fd, rel=1e-5, abs=1e-7 * abs(float(c)))
    g.close()
