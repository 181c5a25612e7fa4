"""The drop-in host driver on the CPU: configuration grammar and byte-exact output files.

The t = 0 output of the driver (setup, boundary fill, WriteData) must equal, byte for byte, the
file the reference itself wrote for the same input (tests/golden/out_blast20x16_t0.npz, made by
tests/golden/make_golden.py with the compiled reference).  No GPU is touched before the first
Advance, so this runs anywhere.
"""
import os
import subprocess

import numpy as np
import pytest

from gaukuk_amd.config import Config
from gaukuk_amd.gkio import GaukukOutput
from tests._driver import driver, write_ic

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "out_blast20x16_t0.npz")


@pytest.mark.parametrize("name", ["cons_00000", "prim_00000"])
def test_output0_bytes_equal_reference(tmp_path, name):
    z = np.load(GOLD, allow_pickle=False)
    write_ic(str(tmp_path / "ic.bin"), z["ic"])
    (tmp_path / "case.in").write_bytes(z["in_" + name].tobytes())
    env = dict(os.environ, GK_IC_FILE=str(tmp_path / "ic.bin"))
    r = subprocess.run([driver("from_file"), "-i", "case.in", "--setup-only"], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(tmp_path / name, np.uint8)
    ref = z[name]
    assert got.size == ref.size
    assert np.array_equal(got, ref)


def test_reader_parses_reference_file(tmp_path):
    z = np.load(GOLD, allow_pickle=False)
    p = tmp_path / "cons_00000"
    z["cons_00000"].tofile(p)
    o = GaukukOutput(str(p))
    assert (o.nvar, o.nx, o.ny, o.nz, o.nGhost) == (5, 20, 16, 1, 2)
    assert o.data.shape == (5, 5, 20, 24)
    assert np.array_equal(o.active(), z["ic"])
    assert o.t == 0.0
    assert not o.data[:, 0].any()          # empty z planes
    assert not o.data[:, 2, 0, 0].any()     # corner ghost


def test_config_grammar():
    text = """
# comment line
   gamma = 1.4   # trailing comment
CFL=0.4
NX 64
bad line without value
tmax = 1e-2
stepmax = -1
"""
    c = Config.from_text(text)
    assert c.get("gamma") == 1.4 and c.get("CFL") == 0.4 and c.get("NX") == 64
    assert c.get("tmax") == 0.01 and c.get("stepmax") == -1
    assert c.get("missing") == 0.0 and c.get("RcOrder", 1) == 1
    assert "bad" not in c.data


def test_driver_rejects_missing_input(tmp_path):
    r = subprocess.run([driver("kh"), "-i", str(tmp_path / "nope.in")], capture_output=True, text=True, timeout=30)
    assert r.returncode != 0 and "Could not open the input file" in r.stderr


def test_driver_usage():
    r = subprocess.run([driver("kh")], capture_output=True, text=True, timeout=30)
    assert r.returncode == 1 and "Usage" in r.stderr
