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


def relerr(got, ref):
    return np.abs(np.asarray(got) - np.asarray(ref)).max() / max(np.abs(ref).max(), 1e-300)


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
    """16x8x8 stand-in for the 64x32x32 differentiation case, FP64, symmetric
    BCs, tight tolerance: x -> DensityFilter (rmin 1.5) -> Compliance, and
    x.grad from torch autograd against central differences of c(x)."""
    nelx, nely, nelz, nl = 16, 8, 8, 3
    g = top3d_amd.Grid(nelx, nely, nelz, nl, bc="symmetric")
    g.set_max_iter(2000)
    vals = fem_ref.splitmix64_densities(nelx * nely * nelz)
    x = g.density(vals)
    F = g.load()
    rmin = 1.5

    def compliance(xx):
        xp = top3d_amd.DensityFilter.apply(xx, g, rmin)
        return top3d_amd.Compliance.apply(xp, g, F, tol=1e-12)

    xr = x.clone().requires_grad_(True)
    c = compliance(xr)
    c.backward()
    grad_x = xr.grad
    idx = g.interior_elements()
    halo = np.setdiff1d(np.arange(g.nelem), idx)
    assert float(grad_x[halo].abs().max()) == 0.0
    # the same gradient as the reference's chain: sensitivity, then the
    # transposed filter (stencil_optimization.c:160-162)
    xphys = g.filter(rmin, x)
    U = g.zeros(0)
    g.solve(xphys, F, U, tol=1e-12)
    _, dc = g.sensitivity(xphys, U)
    ref = g.filter_grad(rmin, dc).cpu().numpy()
    assert relerr(grad_x.cpu().numpy()[idx], ref[idx]) <= 1e-10
    rng = np.random.default_rng(3)
    pick = rng.choice(idx, 8, replace=False)
    h = 1e-4
    for e in pick:
        xp_, xm_ = x.clone(), x.clone()
        xp_[e] += h
        xm_[e] -= h
        fd = (float(compliance(xp_)) - float(compliance(xm_))) / (2 * h)
        assert float(grad_x[e]) == pytest.approx(fd, rel=1e-6, abs=1e-7 * abs(c.item()))
    g.close()


def test_compliance_reports_no_convergence():
    """Identity-row BCs stagnate on heterogeneous densities: the autograd
    compliance raises unless the caller accepts an unconverged solve, and
    then warns and records the iteration count and residual."""
    g = top3d_amd.Grid(16, 8, 8, 3)
    g.set_max_iter(40)
    rng = np.random.default_rng(5)
    x = g.density(rng.uniform(0.1, 1.0, 16 * 8 * 8)).requires_grad_(True)
    F = g.load()
    with pytest.raises(top3d_amd.Top3dError, match="no convergence"):
        top3d_amd.Compliance.apply(x, g, F)
    with pytest.warns(RuntimeWarning, match="CG stopped"):
        c = top3d_amd.Compliance.apply(x, g, F, allow_no_converge=True)
    info = g.last_compliance_solve
    assert not info["converged"] and info["iterations"] == 40 and info["relres"] > 1e-5
    c.backward()
    assert torch.isfinite(x.grad).all()
    g.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_design_loop_autograd_matches_design_loop_c1(precision):
    """c1 (48x24x24, 3 levels, 5 design iterations) with the OC update fed by
    torch autograd (DensityFilter -> Compliance -> backward) follows the
    trajectory of the reference-ordered loop (sensitivity, then the
    transposed filter) to 1e-12, with the same CG counts."""
    ref = json.load(open(os.path.join(GOLDEN, "c1_fp64.json")))
    shape = (ref["nelx"], ref["nely"], ref["nelz"], ref["nl"])
    g1 = top3d_amd.Grid(*shape, precision=precision)
    recs1, st1 = top3d_amd.design_loop(g1, ref["iters"])
    g2 = top3d_amd.Grid(*shape, precision=precision)
    recs2, st2 = top3d_amd.design_loop_autograd(g2, ref["iters"])
    assert len(recs1) == len(recs2) == 5
    for a, b in zip(recs1, recs2):
        assert a["cgiter"] == b["cgiter"], (a, b)
        assert abs(a["c"] - b["c"]) <= 1e-12 * abs(a["c"]), (a, b)
        assert abs(a["vol"] - b["vol"]) <= 1e-12
    idx = g1.interior_elements()
    dx = float((st1.xPhys[idx] - st2.xPhys[idx]).abs().max())
    assert dx <= 1e-12, dx
    for a, b in zip(recs2, ref["records"]):
        assert abs(a["c"] - b["c"]) <= (1e-8 if precision == "fp64" else 1e-5) * abs(b["c"])
    g1.close()
    g2.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_design_loop_c2_matches_oracle_trajectory(precision):
    """128x64x64 cantilever (coarse 16x8x8, 4 levels), 20 design iterations
    against the oracle fixture: per-iteration compliance within 1e-8 (FP64)
    or 1e-5 (mixed) relative, equal CG iteration counts in FP64, final
    filtered densities within 1e-3 (fixture quantized to 1/65535)."""
    ref = json.load(open(os.path.join(GOLDEN, "c2_fp64.json")))
    print(f"[gpu design c2] {precision}", flush=True)
    g = top3d_amd.Grid(ref["nelx"], ref["nely"], ref["nelz"], ref["nl"], precision=precision)
    recs, st = top3d_amd.design_loop(g, ref["iters"])
    assert len(recs) == len(ref["records"])
    tol = 1e-8 if precision == "fp64" else 1e-5
    worst = 0.0
    for a, b in zip(recs, ref["records"]):
        rel = abs(a["c"] - b["c"]) / abs(b["c"])
        worst = max(worst, rel)
        assert rel <= tol, (a, b)
        assert abs(a["vol"] - b["vol"]) <= 1e-6
        if precision == "fp64":
            assert a["cgiter"] == b["cgiter"], (a, b)
    xq = np.load(os.path.join(GOLDEN, "c2_xphys_fp64.npz"))["xphys_u16"]
    got = st.xPhys.cpu().numpy()[g.interior_elements()]
    dx = np.abs(got - xq.astype(np.float64) / 65535.0).max()
    print(f"[gpu design c2] {precision}: max rel compliance diff {worst:.2e}, "
          f"final density max diff {dx:.2e}, cg {[r['cgiter'] + 1 for r in recs]}", flush=True)
    assert dx <= 1e-3
    g.close()


def test_autograd_c3_gradient_fd_and_oracle_sensitivity():
    """Differentiation case: 64x32x32 elements, 3 levels (coarse 16x8x8),
    densities i.i.d. uniform in [0.1, 1] (seed 2024), FP64, symmetric BCs so
    that CG converges to 1e-12 on the heterogeneous field.  The backward's
    element sensitivities equal the oracle's at the same displacement, and
    the gradient through the filter matches central differences on 32
    seeded elements to 1e-6 relative.  The differences use c = F^T u (a few
    dozen loaded DOFs, so c is evaluated to ~1e-16 relative) with h = 1e-4:
    2 h |g| is ~5e-8, so the ~1e-14 relative rounding of the 65k-term
    element-energy sum, or h = 1e-6, would leave errors of a few 1e-6."""
    nelx, nely, nelz, nl = 64, 32, 32, 3
    print("[gpu autograd c3] 64x32x32", flush=True)
    g = top3d_amd.Grid(nelx, nely, nelz, nl, bc="symmetric")
    g.set_max_iter(5000)
    o = pyoracle.Grid(nelx, nely, nelz, nl, bcsym=True)
    rng = np.random.default_rng(2024)
    x = g.density(rng.uniform(0.1, 1.0, nelx * nely * nelz))
    F = g.load()
    rmin = 1.5
    xphys = g.filter(rmin, x)
    xp = xphys.clone().requires_grad_(True)
    c = top3d_amd.Compliance.apply(xp, g, F, tol=1e-12)
    c.backward()
    # the same displacement through the plain solve; oracle sensitivities
    U = g.zeros(0)
    g.solve(xphys, F, U, tol=1e-12)
    c_o, dc_o = o.sensitivity(xphys.cpu().numpy(), U.cpu().numpy())
    idx = g.interior_elements()
    assert abs(c.item() - c_o) <= 1e-10 * abs(c_o)
    assert abs(c.item() - float(torch.dot(F, U))) <= 1e-10 * abs(c_o)
    assert relerr(xp.grad.cpu().numpy()[idx], dc_o[idx]) <= 1e-10
    grad_x = g.filter_grad(rmin, xp.grad).cpu().numpy()

    def compliance(xx):
        Ux = g.zeros(0)
        g.solve(g.filter(rmin, xx), F, Ux, tol=1e-12)
        return float(torch.dot(F, Ux))

    pick = np.random.default_rng(7).choice(idx, 32, replace=False)
    h = 1e-4
    worst = 0.0
    for e in pick:
        xp_, xm_ = x.clone(), x.clone()
        xp_[e] += h
        xm_[e] -= h
        fd = (compliance(xp_) - compliance(xm_)) / (2 * h)
        worst = max(worst, abs(grad_x[e] - fd) / abs(fd))
        assert grad_x[e] == pytest.approx(fd, rel=1e-6), e
    print(f"[gpu autograd c3] worst FD rel diff {worst:.2e}", flush=True)
    g.close()
    o.close()
