"""GPU parity of the state solve with the whole x=0 face clamped and the
MATLAB line load (stencil_grid_utility.c:19-32, top3dmgcg_reference.m:59-62)
against the CPU oracle, through the C ABI.

Tolerances: FP64 form -- the oracle's CG iteration count and the same iterate
to 1e-9 relative; mixed form -- compliance within 1e-5 relative."""
import numpy as np
import pytest

import pyoracle
import top3d_amd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bc", ["reference", "symmetric"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_face_clamp_line_load_solve(precision, bc):
    nelx, nely, nelz, nl = 32, 16, 16, 4
    print(f"[gpu supports] {precision} {bc}", flush=True)
    g = top3d_amd.Grid(nelx, nely, nelz, nl, precision=precision, bc=bc, support="face")
    o = pyoracle.Grid(nelx, nely, nelz, nl, bcsym=bc == "symmetric", support="face")
    for l in range(nl):
        assert np.array_equal(g.fixed_dofs(l), o.fixed(l))
    F = g.line_load()
    Fo = o.line_load()
    assert np.array_equal(F.cpu().numpy(), Fo)
    x = g.density(fill=float(np.float32(0.2)))
    U = g.zeros(0)
    it, res = g.solve(x, F, U)
    Uo = np.zeros(o.ndof(0))
    ito, reso = o.solve(x.cpu().numpy(), Fo, Uo)
    Ug = U.cpu().numpy()
    c, co = Fo @ Ug, Fo @ Uo
    if precision == "fp64":
        assert it == ito
        assert res == pytest.approx(reso, rel=1e-6)
        assert np.linalg.norm(Ug - Uo) <= 1e-9 * np.linalg.norm(Uo)
    else:
        assert res < 1e-5 and abs(it - ito) <= 2
        assert abs(c - co) <= 1e-5 * abs(co)
    g.close()
    o.close()


@pytest.mark.parametrize("profile", [1, 2])
def test_line_load_profiles_solve(profile):
    # end-weighted and sine edge loads (top3dmgcg_reference.m:57-58), FP64
    nelx, nely, nelz, nl = 32, 16, 16, 4
    print(f"[gpu supports] line load profile {profile}", flush=True)
    g = top3d_amd.Grid(nelx, nely, nelz, nl, support="face")
    o = pyoracle.Grid(nelx, nely, nelz, nl, support="face")
    F = g.line_load(profile)
    Fo = o.line_load(profile)
    np.testing.assert_allclose(F.cpu().numpy(), Fo, rtol=1e-15, atol=1e-17)
    x = g.density(fill=float(np.float32(0.2)))
    U = g.zeros(0)
    it, res = g.solve(x, F, U)
    Uo = np.zeros(o.ndof(0))
    ito, reso = o.solve(x.cpu().numpy(), Fo, Uo)
    assert it == ito
    assert np.linalg.norm(U.cpu().numpy() - Uo) <= 1e-9 * np.linalg.norm(Uo)
    g.close()
    o.close()
