"""The node-centric fine and level-1 kernels (one thread per node, the
reference's gather form of stencil_methods.c:7-519), which the library uses
when TOP3D_FINE_KERNEL=node / TOP3D_LEVEL1_KERNEL=node are set, or for level 1
when the transformed sub-cell matrix lacks its fixed pattern: operators,
Jacobi sweeps and a whole solve against the oracle, in a child process
(the library reads the variables once)."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_node_kernels_match_oracle():
    env = dict(os.environ, TOP3D_FINE_KERNEL="node", TOP3D_LEVEL1_KERNEL="node")
    print("[gpu node kernels] child process", flush=True)
    run = subprocess.run([sys.executable, "-u", os.path.join(HERE, "node_kernels_check.py")],
                         env=env, capture_output=True, text=True, timeout=240)
    print(run.stdout[-2000:], run.stderr[-2000:], flush=True)
    assert run.returncode == 0
