"""The top3d command line (top3d.c:14-88) and its design loop top3dmgcg
(stencil_optimization.c:13-223), run as a program on the GPU: c1
(-x 12 -y 6 -z 6 -l 3 -i 5, 48x24x24) against the FP64 oracle fixture.

- Every verbose line (stencil_optimization.c:181-185) has the fixture's CG
  iteration count, compliance to the printed 4 digits, volume and change.
- The summary line (stencil_optimization.c:193) has the fixture's last
  compliance to its 6 printed decimals and the total CG iteration count.
- With -w 1 the density snapshot (writeDensityVtkFile,
  stencil_optimization.c:399-497) holds the fixture's final densities to the
  %e format's 6 digits (FP64) or to 1e-3 (mixed, the final-density bar).
"""
import json
import os
import re
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "topopt-in-openmp-gpu_amd", "bin", "top3d")
GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.mark.parametrize("precision", [0, 1])
def test_cli_c1_matches_oracle_fixture(tmp_path, precision):
    ref = json.load(open(os.path.join(GOLDEN, "c1_fp64.json")))["records"]
    cmd = [BIN, "-x", "12", "-y", "6", "-z", "6", "-l", "3", "-i", "5", "-v", "1", "-w", "1",
           "-p", str(precision)]
    print("[gpu cli] " + " ".join(cmd[1:]), flush=True)
    run = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    lines = [m for m in re.finditer(
        r"It\.:\s*(\d+) Obj\.:(\S+) Vol\.:\s*(\S+) ch\.:(\S+) relres: (\S+) iters:\s*(\d+)",
        run.stdout)]
    assert len(lines) == 5, run.stdout
    tol = 1e-8 if precision == 0 else 1e-5
    for m, r in zip(lines, ref):
        loop, c, vol, ch, relres, cg = m.groups()
        assert int(loop) == r["loop"]
        if precision == 0:
            assert int(cg) == r["cgiter"]
        assert float(c) == pytest.approx(r["c"], rel=max(tol, 5e-4))
        assert float(vol) == pytest.approx(r["vol"], abs=1e-3)
        assert float(ch) == pytest.approx(r["change"], abs=1e-2)
        assert float(relres) < 1e-5
    summary = [ln for ln in run.stdout.splitlines() if re.match(r"^\s*5\s+\d+\.\d{6}\s", ln)]
    assert len(summary) == 1, run.stdout
    f = summary[0].split()
    assert float(f[1]) == pytest.approx(ref[-1]["c"], rel=max(tol, 0), abs=1.5e-6)
    assert int(f[4]) == 48 * 24 * 24
    if precision == 0:
        assert int(f[5]) == sum(r["cgiter"] for r in ref)
    # density snapshot written by -w 1
    vtu = tmp_path / "out_48_24_24.vtu"
    assert vtu.exists() and (tmp_path / "out_48_24_24_halo.vtu").exists()
    root = ET.parse(vtu).getroot()
    dens = [da for da in root.iter("DataArray") if da.get("Name") == "density"][0]
    got = np.array(dens.text.split(), float)
    want = np.load(os.path.join(GOLDEN, "c1_xphys_fp64.npz"))["xphys"]
    assert got.shape == want.shape
    # FP64: the %e format's 6 digits; mixed: the 1e-3 final-density bar
    assert np.abs(got - want).max() <= (1e-5 if precision == 0 else 1e-3)
