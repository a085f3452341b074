"""GPU checks at the benchmark's full size: the 256x128x128 cantilever with 5
levels (12.8 M DOFs), through properties that need no CPU oracle run (the
oracle takes tens of seconds per CG iteration at this size).

- The FP64 fine operator is symmetric on the free DOFs: <K x, y> = <x, K y>
  to 1e-12 relative, for random densities and vectors.
- Both solve forms reach the tolerance on the true residual ||F - K U|| / ||F||
  computed with the FP64 operator, not only on the CG recurrence.
- The mixed form's compliance is within 1e-5 relative of the FP64 form's, with
  the same CG iteration count (within one)."""
import numpy as np
import pytest
import torch

import top3d_amd

pytestmark = pytest.mark.gpu
SHAPE = (256, 128, 128, 5)


@pytest.fixture(scope="module")
def grids():
    print(f"[gpu full size] {SHAPE}", flush=True)
    g64 = top3d_amd.Grid(*SHAPE)
    g32 = top3d_amd.Grid(*SHAPE, precision="mixed")
    yield g64, g32
    g64.close()
    g32.close()


def free_dof_mask(g):
    """True at the DOFs of real nodes that are not fixed."""
    _, wx, wy, wz, ndof = g.padding(0)
    nelx, nely, nelz = g.nelx, g.nely, g.nelz
    node = np.zeros((wx, wz, wy), bool)
    node[1:nelx + 2, 1:nelz + 2, 1:nely + 2] = True
    mask = np.repeat(node.ravel(), 3)
    mask[g.fixed_dofs(0)] = False
    return torch.from_numpy(mask).to(g.device)


def test_fine_operator_symmetric_on_free_dofs(grids):
    g, _ = grids
    rng = np.random.default_rng(3)
    x = g.density(rng.uniform(0.1, 1.0, g.nelx * g.nely * g.nelz))
    free = free_dof_mask(g)
    gen = torch.Generator(device=g.device).manual_seed(4)
    u = torch.randn(g.ndof(0), dtype=torch.float64, device=g.device, generator=gen) * free
    v = torch.randn(g.ndof(0), dtype=torch.float64, device=g.device, generator=gen) * free
    Ku, Kv = g.apply(x, u), g.apply(x, v)
    a, b = float(Ku @ v), float(u @ Kv)
    assert abs(a - b) <= 1e-12 * max(abs(a), abs(b))


def test_full_size_solves_true_residual_and_mixed_compliance(grids):
    g64, g32 = grids
    x = g64.density(fill=float(np.float32(0.2)))
    F = g64.load()
    res = {}
    for name, g in (("fp64", g64), ("mixed", g32)):
        U = g.zeros(0)
        it, rr = g.solve(x, F, U)
        true = float(torch.linalg.norm(F - g64.apply(x, U)) / torch.linalg.norm(F))
        print(f"[gpu full size] {name}: {it} CG iterations, relres {rr:.3e}, true {true:.3e}",
              flush=True)
        assert rr < 1e-5
        assert true < 1e-5 * (1 + 1e-3)
        res[name] = (it, float(F @ U))
    (it64, c64), (it32, c32) = res["fp64"], res["mixed"]
    assert abs(it32 - it64) <= 1
    assert abs(c32 - c64) <= 1e-5 * abs(c64)
