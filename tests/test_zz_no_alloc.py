"""update_solution_noalloc (per-device static scratch, no device allocation)
against update_solution (scratch from hipMalloc) through the C ABI: two
calls in a row, the second on a smaller input, each bit-identical to the
allocating entry point on the same arrays.  Also: an unmapped host array is
rejected instead of staged."""
import ctypes

import numpy as np
import pytest
import torch

import top3d_amd

pytestmark = pytest.mark.gpu


def _case(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.02, 0.98, n)
    dc = -rng.uniform(0.05, 3.0, n)
    dv = rng.uniform(0.5, 1.5, n)
    return [torch.from_numpy(a).cuda() for a in (x, dv, dc)]


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def test_noalloc_oc_matches_allocating_entry_point():
    lib = top3d_amd.library()
    lib.top3d_set_stream(ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    for n, seed, gval in ((300_000, 1, 0.01), (70_001, 2, -0.03)):
        print(f"[gpu noalloc] n {n} seed {seed}", flush=True)
        x, dv, dc = _case(n, seed)
        ref = torch.zeros_like(x)
        got = torch.full_like(x, -1.0)
        torch.cuda.synchronize()
        lib.top3d_clear_error()
        lib.update_solution(_p(x), _p(ref), _p(dv), _p(dc), n, gval)
        top3d_amd.check("update_solution")
        before = top3d_amd.memory_info()["library_bytes"]
        lib.update_solution_noalloc(_p(x), _p(got), _p(dv), _p(dc), n, gval)
        top3d_amd.check("update_solution_noalloc")
        lib.top3d_synchronize()
        assert top3d_amd.memory_info()["library_bytes"] == before
        assert torch.equal(got, ref)
        assert float((ref - x).abs().max()) <= 0.2 + 1e-12


def test_noalloc_oc_rejects_unmapped_host_arrays():
    lib = top3d_amd.library()
    n = 1000
    x = np.full(n, 0.5)
    xnew = np.zeros(n)
    dv = np.ones(n)
    dc = -np.ones(n)
    lib.top3d_clear_error()
    lib.update_solution_noalloc(x.ctypes.data_as(ctypes.c_void_p),
                                xnew.ctypes.data_as(ctypes.c_void_p),
                                dv.ctypes.data_as(ctypes.c_void_p),
                                dc.ctypes.data_as(ctypes.c_void_p), n, 0.0)
    assert lib.top3d_last_error() == 3  # TOP3D_ERR_UNMAPPED
    lib.top3d_clear_error()
    assert np.all(xnew == 0.0)
