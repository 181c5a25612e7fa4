"""Loading helpers for the golden fixtures in tests/golden/ (made by make_golden.py)."""
import glob
import json
import os

import numpy as np

from gaukuk_amd.problems import Problem

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def run_cases():
    return sorted(os.path.basename(p)[4:-4] for p in glob.glob(os.path.join(GOLDEN, "run_*.npz")))


def load_run(name):
    z = np.load(os.path.join(GOLDEN, f"run_{name}.npz"), allow_pickle=False)
    m = json.loads(str(z["meta"]))
    prob = Problem(
        name=name, nx=int(m["NX"]), ny=int(m["NY"]),
        extent=(m["xmin"], m["xmax"], m["ymin"], m["ymax"], m["zmin"], m["zmax"]),
        gamma=m["gamma"], cfl=m["CFL"], integrator=int(m["integrator"]), rc_order=int(m["RcOrder"]),
        stepmax=int(m["stepmax"]), tmax=m["tmax"], dtoutput=m["dtoutput"],
        bc=(int(m["xleft"]), int(m["xright"]), int(m["yleft"]), int(m["yright"])),
        nz=int(m.get("NZ", 1)), bcz=(int(m.get("zleft", 1)), int(m.get("zright", 1))),
        g=(m["gx"], m["gy"], m["gz"]), ng=int(m["NGhost"]), cons=z["ic"].copy(),
        point=tuple(tuple(m[f"pg{q}_{k}"] for k in ("gm", "x", "y", "z", "vx", "vy", "vz", "rs"))
                    for q in range(int(m.get("npoint", 0)))),
        fixed=_fixed(z, m),
        self_faces=tuple((f, tuple(m[f"sd{f}_{q}"] for q in range(5))) for f in range(4)
                         if int(m.get(f"selfbd_{f}", 0))),
        fix_inner=bool(int(m.get("fixinner", 0))),
        shell=tuple(m[f"shell_{k}"] for k in ("x", "y", "z", "r")) if int(m.get("immersed", 0)) else ())
    return prob, z["final"].copy(), float(z["t"]), int(z["steps"])


def _fixed(z, m):
    kind = int(m.get("fixedbd", 0))
    if kind == 1:
        return tuple(m[f"fix_{q}"] for q in range(5))
    if kind in (2, 3):
        return z["fixed_full"].copy()
    return ()


def load_vec(name):
    return np.load(os.path.join(GOLDEN, f"vec_{name}.npz"), allow_pickle=False)
