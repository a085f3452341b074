"""GPU checks of the measurement hooks that bench.py reports from.

- top3d_time_kernel: every kernel, timed back to back and with the
  memory-side cache flushed before each launch (flag 16), returns a positive
  duration and the same algorithmic bytes in both modes; the level-1 kernels
  run on the moduli transform the timer forms before its launches.
- Timing leaves the state usable: a solve afterwards keeps the CG count and
  compliance of the solve before.
- top3d_sweep_timing counts the fine sweeps of the solves it brackets (7 per
  V-cycle with V(4,4): 3 after the first pre-sweep, 4 post-sweeps).
- An unknown kernel index is rejected with TOP3D_ERR_ARG."""
import ctypes

import pytest
import torch

import top3d_amd

pytestmark = pytest.mark.gpu
SHAPE = (32, 16, 16, 3)


@pytest.fixture(scope="module")
def grid():
    print(f"[gpu timing] {SHAPE}", flush=True)
    g = top3d_amd.Grid(*SHAPE, precision="mixed")
    yield g
    g.close()


def test_time_kernel_warm_and_cold(grid):
    st = top3d_amd.DesignState(grid)
    top3d_amd.design_step(st)
    torch.cuda.synchronize()
    for name in top3d_amd.Grid.KERNELS:
        ms, by = grid.time_kernel(name, 3)
        cms, cby = grid.time_kernel(name, 3, cold=True)
        print(f"[gpu timing] {name}: {ms * 1e3:.1f} us, cold {cms * 1e3:.1f} us", flush=True)
        assert ms > 0 and cms > 0
        assert by > 0 and cby == by


def test_state_usable_after_timing(grid):
    # the same density twice: the design loop's first step, solved again
    # after the timer ran on the solver's buffers
    st1 = top3d_amd.DesignState(grid)
    r1 = top3d_amd.design_step(st1)
    for name in ("level1_sweep_f32", "fine_apply_dot_f64"):
        grid.time_kernel(name, 2, cold=True)
    st2 = top3d_amd.DesignState(grid)
    r2 = top3d_amd.design_step(st2)
    assert r2["cgiter"] == r1["cgiter"]
    assert r2["c"] == pytest.approx(r1["c"], rel=1e-12)


def test_sweep_timing_counts_fine_sweeps(grid):
    st = top3d_amd.DesignState(grid)
    top3d_amd.design_step(st)
    torch.cuda.synchronize()
    grid.sweep_timing(True)
    rec = top3d_amd.design_step(st)
    ms, n, by = grid.sweep_timing(False)
    vcycles = rec["cgiter"] + 1
    print(f"[gpu timing] {n} sweeps in {vcycles}+ V-cycles, {ms * 1e3:.1f} us each", flush=True)
    assert ms > 0 and by > 0
    # 7 per V-cycle; a V-cycle queued after the last iteration may add 7
    assert n in (7 * vcycles, 7 * (vcycles + 1))


def test_time_kernel_rejects_unknown_kernel(grid):
    lib = grid.lib
    ms, by = ctypes.c_double(), ctypes.c_double()
    rc = lib.top3d_time_kernel(ctypes.byref(grid.data), 6, 1, ctypes.byref(ms), ctypes.byref(by))
    assert rc == 2  # TOP3D_ERR_ARG
    lib.top3d_clear_error()
