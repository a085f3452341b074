"""GPU checks at the largest single-GPU size: the 512x256x256 cantilever
(coarse 16x8x8, 6 levels, 101.6 M DOFs), through properties that need no
CPU oracle run.

- Both solve forms reach rtol 1e-5 on the true residual ||F - K U|| / ||F||
  computed with the FP64 fine operator; the mixed form's compliance is
  within 1e-5 relative of the FP64 form's, with the same CG iteration count
  (within one).
- Two design iterations whose OC update is fed by torch autograd
  (DensityFilter -> Compliance -> backward) follow the reference-ordered
  loop (sensitivity, then the transposed filter) to 1e-12.
- The autograd gradient of c(x) through the density filter matches a
  central difference along a random direction of all 33.5 M elements (a
  single-element difference is below FP64 resolution at this size), to
  1e-6 relative.  This check runs with the symmetric BC convention, whose
  V-cycle is symmetric, so that CG reaches 1e-11 on the perturbed densities.
- Every 32-bit offset the tile kernels use stays below 2^31 here:
  3 * nodes(level 0) = 103.6 M.
"""
import numpy as np
import pytest
import torch

import top3d_amd

pytestmark = pytest.mark.gpu
SHAPE = (512, 256, 256, 6)


def true_residual(g64, x, F, U):
    return float(torch.linalg.norm(F - g64.apply(x, U)) / torch.linalg.norm(F))


def test_c5_solves_true_residual_and_mixed_compliance():
    print(f"[gpu c5] {SHAPE} solve forms", flush=True)
    top3d_amd.memory_info(reset_peak=True)
    g64 = top3d_amd.Grid(*SHAPE)
    _, wx, wy, wz, ndof = g64.padding(0)
    assert ndof < 2 ** 31
    x = g64.density(fill=float(np.float32(0.2)))
    F = g64.load()
    res = {}
    for name in ("fp64", "mixed"):
        g = g64 if name == "fp64" else top3d_amd.Grid(*SHAPE, precision="mixed")
        U = g.zeros(0)
        it, rr = g.solve(x, F, U)
        st = g.stats()
        true = true_residual(g64, x, F, U)
        print(f"[gpu c5] {name}: {it + 1} CG iterations, relres {rr:.3e}, true {true:.3e}, "
              f"solve {st['solve_ms'] + st['setup_ms']:.1f} ms", flush=True)
        assert rr < 1e-5
        assert true < 1e-5 * (1 + 1e-3)
        res[name] = (it, float(F @ U))
        if g is not g64:
            g.close()
    (it64, c64), (it32, c32) = res["fp64"], res["mixed"]
    assert abs(it32 - it64) <= 1
    assert abs(c32 - c64) <= 1e-5 * abs(c64)
    mem = top3d_amd.memory_info()
    print(f"[gpu c5] library peak {mem['library_peak_bytes'] / 2**30:.1f} GiB, device used "
          f"{mem['device_used_bytes'] / 2**30:.1f} GiB", flush=True)
    assert 0 < mem["library_peak_bytes"] <= mem["device_total_bytes"]
    g64.close()


def test_c5_autograd_design_steps_match():
    print("[gpu c5] autograd design steps", flush=True)
    recs = {}
    for kind in ("plain", "autograd"):
        g = top3d_amd.Grid(*SHAPE, precision="mixed")
        st = top3d_amd.DesignState(g)
        step = top3d_amd.design_step if kind == "plain" else top3d_amd.design_step_autograd
        recs[kind] = [step(st) for _ in range(2)]
        recs[kind + "_x"] = st.x.clone()
        g.close()
        print(f"[gpu c5] {kind}: " + ", ".join(f"c {r['c']:.10e} cg {r['cgiter'] + 1}"
                                              for r in recs[kind]), flush=True)
    for a, b in zip(recs["plain"], recs["autograd"]):
        assert a["cgiter"] == b["cgiter"]
        assert abs(a["c"] - b["c"]) <= 1e-12 * abs(a["c"])
        assert abs(a["change"] - b["change"]) <= 1e-6
    assert float((recs["plain_x"] - recs["autograd_x"]).abs().max()) <= 1e-12


def test_c5_autograd_gradient_directional_fd():
    """grad(c) . v from autograd against (c(x + h v) - c(x - h v)) / 2h.
    The differences take c = F^T U from cold-started solves: CG from zero
    gives c - c_k = ||e_k||_K^2, quadratic in the error, so c is resolved
    far below the 1e-6 relative size of c(x + h v) - c(x - h v)."""
    print("[gpu c5] directional FD of the autograd gradient", flush=True)
    g = top3d_amd.Grid(*SHAPE, bc="symmetric")
    g.set_max_iter(600)
    rmin, tol, h = 1.5, 1e-11, 3e-4
    idx = g._interior_index()
    gen = torch.Generator(device=g.device).manual_seed(5)
    x = torch.zeros(g.nelem, dtype=torch.float64, device=g.device)
    x[idx] = 0.3 + 0.2 * torch.rand(idx.numel(), dtype=torch.float64, device=g.device,
                                    generator=gen)
    v = torch.zeros_like(x)
    v[idx] = torch.randn(idx.numel(), dtype=torch.float64, device=g.device, generator=gen)
    F = g.load()
    xl = x.clone().requires_grad_(True)
    c = top3d_amd.Compliance.apply(top3d_amd.DensityFilter.apply(xl, g, rmin), g, F, tol=tol)
    c.backward()
    print(f"[gpu c5] c = {float(c.detach()):.12e}, CG {g.last_compliance_solve['iterations']}",
          flush=True)

    def c_cold(xx):
        U = g.zeros(0)
        it, _ = g.solve(g.filter(rmin, xx), F, U, tol=tol)
        return float(F @ U), it

    (cp, itp), (cm, itm) = c_cold(x + h * v), c_cold(x - h * v)
    fd = (cp - cm) / (2 * h)
    an = float(xl.grad @ v)
    print(f"[gpu c5] grad.v {an:.12e}, central difference {fd:.12e} (CG {itp + 1}, {itm + 1}), "
          f"rel {abs(an - fd) / abs(fd):.2e}", flush=True)
    assert an == pytest.approx(fd, rel=1e-6)
    g.close()
