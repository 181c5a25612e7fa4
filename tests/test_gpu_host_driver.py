"""End-to-end drop-in path on the GPU: host driver -> C ABI -> HIP stage kernels.

The driver runs golden cases through its setup_from_file plug-in and the shipped Sod setup
through input/shock_tube.in; final files are compared with the reference's own results at the
FP64 parity bar (1e-12 relative L1 per variable), with identical step counts.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from gaukuk_amd.gkio import GaukukOutput
from tests._driver import HOST, driver, run_from_file
from tests._fixtures import load_run

pytestmark = pytest.mark.gpu


def rel_l1(a, b):
    scale = np.sum(np.abs(b[0]))
    return np.array([np.sum(np.abs(a[n] - b[n])) / (np.sum(np.abs(b[n])) or scale) for n in range(5)])


@pytest.mark.parametrize("name", ["kh64_rk3", "rt32x96_rk2_g", "blast48x56_mixed", "sod128_fe_rc1",
                                  "point48_rk2_bc3", "binary40x44_rk3_g", "fixed48x40_rk3",
                                  "fixedvar52x44_rk2_g"])
def test_driver_golden(tmp_path, name):
    prob, final, t, steps = load_run(name)
    print(f"driver case {name}", flush=True)
    out, total = run_from_file(str(tmp_path), prob, extra_args=("--quiet",))
    assert total == steps
    files = sorted(f for f in os.listdir(tmp_path) if f.startswith("cons_"))
    o = GaukukOutput(str(tmp_path / files[-1]))
    assert abs(o.t - t) <= 1e-12 * t
    err = rel_l1(o.active(), final)
    print(name, err)
    assert np.all(err <= 1e-12), err


def test_driver_sod_setup_matches_reference(tmp_path):
    prob, final, t, steps = load_run("sod256")
    shutil.copy(os.path.join(HOST, "input", "shock_tube.in"), tmp_path / "shock_tube.in")
    r = subprocess.run([driver("shock_tube"), "-i", "shock_tube.in"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "step = 280," in r.stdout and "total steps =    280" in r.stdout
    o = GaukukOutput(str(tmp_path / "prim_00001"))
    assert o.t == pytest.approx(0.2, abs=1e-12)
    from oracle import gko
    ref_prim = gko.cons_to_prim(final.reshape(5, -1), 1.4).reshape(5, 1, 256)
    err = rel_l1(o.active(), ref_prim)
    assert np.all(err[[0, 1, 4]] <= 1e-12), err
