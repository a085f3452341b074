"""Checks the node-centric fine and level-1 operator kernels against the
CPU oracle.  Run as a separate process with TOP3D_FINE_KERNEL=node and
TOP3D_LEVEL1_KERNEL=node in the environment (the library reads them once);
tests/test_gpu_node_kernels.py does so.  Exits non-zero on a mismatch."""
import ctypes
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "topopt-in-openmp-gpu_amd")):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402
import torch  # noqa: E402

import pyoracle  # noqa: E402
import top3d_amd  # noqa: E402


def relerr(got, ref):
    return np.abs(np.asarray(got) - np.asarray(ref)).max() / max(np.abs(ref).max(), 1e-300)


def cuda(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0", dtype)


def main():
    assert os.environ.get("TOP3D_FINE_KERNEL") == "node"
    assert os.environ.get("TOP3D_LEVEL1_KERNEL") == "node"
    worst = {}
    for nelx, nely, nelz, nl in ((16, 8, 8, 3), (24, 16, 8, 4)):
        g = top3d_amd.Grid(nelx, nely, nelz, nl)
        o = pyoracle.Grid(nelx, nely, nelz, nl)
        rng = np.random.default_rng(nelx)
        xg = g.density(rng.uniform(0.1, 1.0, nelx * nely * nelz))
        xo = xg.cpu().numpy()
        v = rng.standard_normal(g.ndof(0))
        worst["fine fp64"] = max(worst.get("fine fp64", 0),
                                 relerr(g.apply(xg, cuda(v)).cpu().numpy(), o.matvec0(xo, v)))
        got = g.apply_f32(xg, cuda(v, torch.float32)).cpu().numpy()
        worst["fine fp32"] = max(worst.get("fine fp32", 0), relerr(
            got, o.matvec0(xo, v.astype(np.float32).astype(np.float64))))
        v1 = rng.standard_normal(g.ndof(1))
        worst["level1 fp64"] = max(worst.get("level1 fp64", 0), relerr(
            g.apply_level(1, xg, cuda(v1)).cpu().numpy(), o.matvec_l(1, xo, v1)))
        for l in (0, 1):
            inv = o.invdiag(xo, l)
            u = rng.standard_normal(g.ndof(l))
            b = rng.standard_normal(g.ndof(l))
            ug = g.jacobi(xg, 3, cuda(inv), cuda(u), cuda(b), l=l).cpu().numpy()
            uo, tmp = u.copy(), np.zeros_like(u)
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
            key = f"jacobi l{l}"
            worst[key] = max(worst.get(key, 0), relerr(ug, uo))
        # a whole solve with both node kernels in the V-cycle
        F = g.load()
        xu = g.density(fill=float(np.float32(0.2)))
        U = g.zeros(0)
        it, _ = g.solve(xu, F, U)
        Uo = np.zeros(o.ndof(0))
        ito, _ = o.solve(xu.cpu().numpy(), F.cpu().numpy(), Uo)
        assert it == ito, (it, ito)
        worst["solve U"] = max(worst.get("solve U", 0),
                               np.linalg.norm(U.cpu().numpy() - Uo) / np.linalg.norm(Uo))
        g.close()
        o.close()
    limits = {"fine fp64": 1e-13, "fine fp32": 2e-6, "level1 fp64": 1e-13, "jacobi l0": 1e-12,
              "jacobi l1": 1e-12, "solve U": 1e-9}
    print(" ".join(f"{k} {v:.2e}" for k, v in worst.items()), flush=True)
    bad = [k for k, v in worst.items() if not v <= limits[k]]
    if bad:
        print("node kernels: mismatch in " + ", ".join(bad), flush=True)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
