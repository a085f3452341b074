"""Writes regression fixtures of the CPU oracle (oracle/top3d_oracle.c).

The reference ships no golden vectors and cannot be built in this image, so
these fixtures are produced by the oracle itself; they let the GPU tests
compare against oracle runs too long to repeat inside a test:
  c1_fp64.json / c1_asship.json   48x24x24, 3 levels, 5 design iterations
  c2_fp64.json                    128x64x64, 4 levels, 20 design iterations
  c1_xphys_fp64.npz               final filtered densities of c1 (interior)
  c2_xphys_fp64.npz               final filtered densities of c2 (interior),
                                  quantized to uint16 steps of 1/65535
Run from the repository root: python tests/golden/make_golden.py [c1|c2]
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import pyoracle  # noqa: E402

CASES = {
    "c1": dict(nelx=48, nely=24, nelz=24, nl=3, iters=5, typings=("fp64", "asship")),
    "c2": dict(nelx=128, nely=64, nelz=64, nl=4, iters=20, typings=("fp64",)),
}


def interior(x, nelx, nely, nelz):
    wy, wz = nely + 3, nelz + 3
    ey, ez = wy - 1, wz - 1
    i, k, j = np.meshgrid(np.arange(1, nelx + 1), np.arange(1, nelz + 1),
                          np.arange(1, nely + 1), indexing="ij")
    return x[(i * ey * ez + k * ey + j).ravel()]


def run(name):
    c = CASES[name]
    for typing in c["typings"]:
        t0 = time.time()
        recs, xphys, _ = pyoracle.top3d(c["nelx"], c["nely"], c["nelz"], c["nl"], c["iters"],
                                        typing=typing)
        out = dict(case=name, typing=typing, nelx=c["nelx"], nely=c["nely"], nelz=c["nelz"],
                   nl=c["nl"], iters=c["iters"], volfrac=0.2, rmin=1.5, cgtol=1e-5,
                   records=recs, seconds=time.time() - t0,
                   generator="oracle/top3d_oracle.c via tests/golden/make_golden.py")
        with open(os.path.join(HERE, f"{name}_{typing}.json"), "w") as f:
            json.dump(out, f, indent=1)
        if name == "c1" and typing == "fp64":
            np.savez_compressed(os.path.join(HERE, "c1_xphys_fp64.npz"),
                                xphys=interior(xphys, c["nelx"], c["nely"], c["nelz"]))
        if name == "c2" and typing == "fp64":
            xi = interior(xphys, c["nelx"], c["nely"], c["nelz"])
            q = np.rint(np.clip(xi, 0.0, 1.0) * 65535.0).astype(np.uint16)
            np.savez_compressed(os.path.join(HERE, "c2_xphys_fp64.npz"), xphys_u16=q)
        print(name, typing, f"{time.time() - t0:.1f}s", flush=True)


if __name__ == "__main__":
    for n in (sys.argv[1:] or list(CASES)):
        run(n)
