"""Drop-in parity of the reference's own setup files on the GPU.

Each setup file of the reference (src/setup/setup_XXX.cpp) is compiled unchanged against this host
driver (`make -C gaukuk_amd/host external SETUP_SRC=...`, done by the build where the reference
sources are present) and run on the reference's shipped input file at a reduced size (2D, and the
3D inputs blast.in and shock_tube.in also in 3D).  The
fixtures (tests/golden/setup_XXX.npz) come from the reference built with the same setup file and
the same input values; results must agree at the FP64 parity bar (1e-12 relative L1 per variable)
with identical step counts.  The setups exercise periodic, reflective, outflow and fixed faces,
constant, point and binary gravity, the inner fixed region, a self-defined inflow face and the
immersed shell.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from gaukuk_amd.config import write_in
from gaukuk_amd.gkio import GaukukOutput
from tests._driver import HOST
from tests._fixtures import GOLDEN

pytestmark = pytest.mark.gpu

SETUPS = ["blast", "disk", "disk_planet", "kh", "rt", "shock_tube", "solid_body", "wave_test", "blast3d",
          "shock_tube3d",
          # this project's tests/golden/setup_statebd.cpp in the same plug-in format: self-defined
          # faces whose ghost values depend on the state on all four sides (the boundary hook)
          "statebd",
          # tests/golden/setup_statebd3d.cpp: the same on all six sides of a 3D grid, then under a
          # point mass in 3D and in 2D
          "statebd3d", "statebd3d_pg", "statebd_pg"]


def rel_l1(a, b):
    scale = np.sum(np.abs(b[0]))
    return np.array([np.sum(np.abs(a[n] - b[n])) / (np.sum(np.abs(b[n])) or scale) for n in range(5)])


@pytest.mark.parametrize("name", SETUPS)
def test_reference_setup_file(tmp_path, name):
    z = np.load(os.path.join(GOLDEN, f"setup_{name}.npz"), allow_pickle=False)
    binary = os.path.join(HOST, "bin", f"gaukuk_external_{str(z['setup'])}.sim")
    if not os.path.exists(binary):
        pytest.skip(f"{binary} not built (the reference setup files were not available at build time)")
    vals = json.loads(str(z["meta"]))
    kind = str(z["kind"])
    write_in(str(tmp_path / "case.in"), vals)
    print(f"reference setup {name} {int(vals['NX'])}x{int(vals['NY'])}", flush=True)
    r = subprocess.run([binary, "-i", "case.in", "--quiet"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    assert f"total steps = {int(z['steps']):6d}" in r.stdout
    files = sorted(f for f in os.listdir(tmp_path) if f.startswith(kind + "_"))
    o = GaukukOutput(str(tmp_path / files[-1]))
    t = float(z["t"])
    assert abs(o.t - t) <= 1e-12 * t
    err = rel_l1(o.active(), z["final"])
    print(name, err)
    assert np.all(err <= 1e-12), err
