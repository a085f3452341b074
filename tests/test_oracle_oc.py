"""Pins the oracle's design-loop updates: or_oc (update_solution,
gpu_definitions.c:127-157), or_volume (compute_volume, :94-110) and
or_change (compute_change, :112-125).

The OC update is checked through the optimality conditions it must satisfy,
not by restating its loop:
  * every element sits in [0, 1] and within the move limit 0.2 of x;
  * the elements strictly inside both bounds satisfy the OC fixed point
    (xnew / x)^2 = -dc / (dv * lambda) with ONE multiplier lambda for all of
    them (this fixes the exponent 1/2);
  * an element on a bound sits there because the unconstrained update for
    that lambda lies beyond it (this fixes the move limit and the clamps);
  * lambda balances the volume: g + sum(dv * (xnew - x)) vanishes to the
    resolution of a bisection stopped at (l2 - l1) / (l1 + l2) <= 1e-6
    (this fixes the bisection direction and the volume term);
and a uniform case has the closed form xnew = x - g / (n * dv).
Copies of the oracle built with a deliberately wrong move limit, exponent,
bisection direction or a dropped volume term must each fail these checks.
"""
import os
import subprocess

import numpy as np
import pytest

import pyoracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOVE = 0.2


def oc_violations(x, dv, dc, g, xnew):
    """Reasons why xnew is not the OC update of (x, dv, dc, g); empty if it is."""
    bad = []
    eps = 1e-12
    lo = np.maximum(0.0, x - MOVE)
    hi = np.minimum(1.0, x + MOVE)
    if (xnew < lo - eps).any() or (xnew > hi + eps).any():
        bad.append("outside [max(0, x-0.2), min(1, x+0.2)]")
        return bad
    inner = (xnew > lo + 1e-9) & (xnew < hi - 1e-9)
    if inner.sum() < 10:
        bad.append(f"only {inner.sum()} unclamped elements")
        return bad
    lam_e = -dc[inner] * x[inner] ** 2 / (dv[inner] * xnew[inner] ** 2)
    lam = np.median(lam_e)
    spread = np.abs(lam_e / lam - 1).max()
    if spread > 1e-10:
        bad.append(f"unclamped elements disagree on the multiplier (spread {spread:.2e})")
    B = x * np.sqrt(-dc / (dv * lam))  # unconstrained update for that multiplier
    at_hi = ~inner & (np.abs(xnew - hi) <= 1e-9)
    at_lo = ~inner & (np.abs(xnew - lo) <= 1e-9)
    if (~inner & ~at_hi & ~at_lo).any():
        bad.append("elements neither free nor on a bound")
    if (B[at_hi] < hi[at_hi] * (1 - 1e-9)).any():
        bad.append("elements clamped from above whose update lies below the bound")
    if (B[at_lo] > lo[at_lo] * (1 + 1e-9) + 1e-12).any():
        bad.append("elements clamped from below whose update lies above the bound")
    # volume balance: a relative bracket of 1e-6 on lambda moves each free
    # element by at most 0.5e-6 of its value
    resid = g + np.sum(dv * (xnew - x))
    tol = 1e-6 * np.sum(dv[inner] * xnew[inner]) + 1e-12
    if abs(resid) > tol:
        bad.append(f"volume residual {resid:.3e} above {tol:.3e}")
    return bad


def random_case(seed, n=4000):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.02, 0.98, n)
    dc = -rng.uniform(0.05, 3.0, n) * rng.choice([1.0, 1e-3, 30.0], n, p=[0.8, 0.1, 0.1])
    dv = rng.uniform(0.5, 1.5, n)
    return x, dc, dv


@pytest.fixture(scope="module")
def grid():
    g = pyoracle.Grid(8, 4, 4, 3)
    yield g
    g.close()


@pytest.mark.parametrize("seed,gval", [(1, 0.01), (2, -0.02), (3, 0.3), (4, 0.0)])
def test_oc_satisfies_optimality_conditions(grid, seed, gval):
    x, dc, dv = random_case(seed)
    xnew = grid.oc(x, dv, dc, gval)
    assert oc_violations(x, dv, dc, gval, xnew) == []


def test_oc_uniform_closed_form(grid):
    n, x0, a, b, gval = 5000, 0.4, 1.7, 0.9, 7.5
    x = np.full(n, x0)
    xnew = grid.oc(x, np.full(n, b), np.full(n, -a), gval)
    expect = x0 - gval / (n * b)  # g + n b (xnew - x0) = 0
    assert np.allclose(xnew, expect, rtol=1e-6, atol=0)
    assert np.ptp(xnew) == 0.0


def test_oc_clamps_at_move_limit_and_bounds(grid):
    # huge, tiny and zero sensitivities push elements onto every bound
    x = np.array([0.5, 0.5, 0.1, 0.95, 0.5] * 200)
    dc = -np.array([1e6, 1e-9, 0.0, 1e6, 1.0] * 200)
    dv = np.ones_like(x)
    xnew = grid.oc(x, dv, dc, 0.0)
    assert np.allclose(xnew[0::5], 0.7) and np.allclose(xnew[1::5], 0.3)
    assert np.all(xnew[2::5] == 0.0) and np.all(xnew[3::5] == 1.0)


def _mutant(tmp_path_factory, flags):
    out = tmp_path_factory.mktemp("oc_mutant") / ("lib" + "_".join(
        f.strip("-D").replace("=", "").replace(".", "") for f in flags) + ".so")
    subprocess.run(["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", *flags,
                    os.path.join(ROOT, "oracle", "top3d_oracle.c"), "-o", str(out), "-lm"],
                   check=True, capture_output=True)
    return str(out)


@pytest.mark.parametrize("flags", [["-DOR_OC_MOVE=0.1"], ["-DOR_OC_ETA=1.0"],
                                   ["-DOR_OC_ETA=0.25"], ["-DOR_OC_FLIP"],
                                   ["-DOR_OC_DROP_G"]],
                         ids=["move0.1", "eta1", "eta0.25", "flipped", "no_g"])
def test_oc_checks_reject_wrong_updates(tmp_path_factory, flags):
    """Each deliberately wrong oracle build fails the checks above on at
    least one of the cases."""
    lib = _mutant(tmp_path_factory, flags)
    g = pyoracle.Grid(8, 4, 4, 3, typing=lib)
    failures = 0
    for seed, gval in [(1, 0.01), (3, 0.3)]:
        x, dc, dv = random_case(seed)
        if oc_violations(x, dv, dc, gval, g.oc(x, dv, dc, gval)):
            failures += 1
    n, x0, a, b, gval = 5000, 0.4, 1.7, 0.9, 7.5
    xu = g.oc(np.full(n, x0), np.full(n, b), np.full(n, -a), gval)
    if not np.allclose(xu, x0 - gval / (n * b), rtol=1e-6, atol=0):
        failures += 1
    g.close()
    assert failures > 0


def interior_index(nelx, nely, nelz, wy, wz):
    i, k, j = np.meshgrid(np.arange(1, nelx + 1), np.arange(1, nelz + 1),
                          np.arange(1, nely + 1), indexing="ij")
    return (i * (wy - 1) * (wz - 1) + k * (wy - 1) + j).ravel()


def test_volume_is_the_interior_sum():
    nelx, nely, nelz = 12, 8, 4
    g = pyoracle.Grid(nelx, nely, nelz, 3)
    _, wx, wy, wz, _ = g.padding(0)
    rng = np.random.default_rng(11)
    x = np.full(g.nelem(), 1e6)  # halo elements must not count
    idx = interior_index(nelx, nely, nelz, wy, wz)
    vals = rng.uniform(0, 1, idx.size)
    x[idx] = vals
    assert g.volume(x) == pytest.approx(vals.sum(), rel=1e-14)
    g.close()


def test_change_is_max_abs_difference_and_copies(grid):
    rng = np.random.default_rng(12)
    n = grid.nelem()
    x = rng.uniform(0, 1, n)
    xnew = rng.uniform(0, 1, n)
    ch, xafter = grid.change(x, xnew)
    assert ch == np.float32(np.abs(x - xnew).max())
    assert np.array_equal(xafter, xnew)
