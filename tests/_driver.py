"""Helpers to run the host driver binaries (gaukuk_amd/host/bin/gaukuk_<setup>.sim)."""
import os
import re
import subprocess

import numpy as np

from gaukuk_amd.config import write_in

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
HOST = os.path.join(ROOT, "gaukuk_amd", "host")


def driver(setup):
    path = os.path.join(HOST, "bin", f"gaukuk_{setup}.sim")
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} missing: run __graft_entry__.build() or `make -C gaukuk_amd/host drivers`")
    return path


def write_ic(path, cons):
    with open(path, "wb") as f:
        np.array([5, cons.shape[-1], cons.shape[-2]], np.int32).tofile(f)
        np.ascontiguousarray(cons, np.float64).tofile(f)


def run_from_file(workdir, prob, extra_args=(), in_values=None, timeout=600):
    """Run gaukuk_from_file.sim on `prob` in workdir; returns (stdout, total steps)."""
    ic = os.path.join(workdir, "ic.bin")
    write_ic(ic, prob.cons)
    inp = os.path.join(workdir, "case.in")
    write_in(inp, in_values if in_values is not None else prob.in_values())
    env = dict(os.environ, GK_IC_FILE=ic)
    if np.ndim(prob.fixed) == 3:
        env["GK_FIXED_FILE"] = os.path.join(workdir, "fixed.bin")
        write_ic(env["GK_FIXED_FILE"], prob.fixed)
    elif np.ndim(prob.fixed) == 4:
        env["GK_FIXED_FILE"] = os.path.join(workdir, "fixed.bin")
        with open(env["GK_FIXED_FILE"], "wb") as f:
            np.array([5, prob.fixed.shape[3], prob.fixed.shape[2], prob.fixed.shape[1]], np.int32).tofile(f)
            np.ascontiguousarray(prob.fixed, np.float64).tofile(f)
    r = subprocess.run([driver("from_file"), "-i", inp, *extra_args], cwd=workdir, env=env,
                       capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"driver failed ({r.returncode}): {r.stderr[-2000:]}")
    m = re.search(r"total steps =\s*(\d+)", r.stdout)
    return r.stdout, int(m.group(1)) if m else None
