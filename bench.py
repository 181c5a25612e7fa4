#!/usr/bin/env python3
"""Benchmark of the per-timestep hydro update on one MI355X.

Workload: Kelvin-Helmholtz 2048x2048, FP64, PLM + HLLC + RK3, CFL 0.4, periodic, seeded ICs.
A step is one full RK3 step (3 fused stage kernels + 2 one-thread clock kernels) over the
whole grid with the state resident in HBM.  Prints one JSON line:
  value         whole-job zone-updates/s over exactly --steps timed steps
  roofline      the stage kernel with the most total time, its algorithmic bytes per launch
                divided by its average launch duration from HIP events on the library's stream
  cpu_baseline  the oracle restatement of the reference's OpenMP CPU path (fast flags) timed on
                this host's cores over a bounded sample of the same workload
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "zone-updates/s & %HBM roofline; time-to-L1-error learned/FP32 vs plain FP64"
HBM_PEAK_GBS = 8000.0
NX = NY = 2048
# The three RK3 stage kernels of a step, with the algorithmic HBM bytes per cell of one launch
# (FP64): the first stage reads U^n and writes U1 (80 B); the two combine stages read U_s and
# U^n and write one state (120 B).
STAGES = [("stage_kernel<double, VanLeer, THREE=false, GRAV=false, LAST=false> (RK3 stage 1)", 80.0),
          ("stage_kernel<double, VanLeer, THREE=true, GRAV=false, LAST=false> (RK3 stage 2)", 120.0),
          ("stage_kernel<double, VanLeer, THREE=true, GRAV=false, LAST=true> (RK3 stage 3)", 120.0)]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(prob, budget_s: float = 12.0):
    """Time the oracle (fast build) on this host: a bounded sample of the same KH workload."""
    threads = int(os.environ.get("GK_CPU_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    from oracle import gko  # test infrastructure: the CPU baseline only
    # one probe step sizes the sample to about budget_s seconds of stepping
    s1 = gko.OracleSim(prob.with_(stepmax=1), flavour="fast")
    t0 = time.perf_counter()
    s1.advance(1e30)
    one = time.perf_counter() - t0
    del s1
    n = max(2, min(200, int(budget_s / max(one, 1e-3))))
    s2 = gko.OracleSim(prob.with_(stepmax=n), flavour="fast")
    t0 = time.perf_counter()
    s2.advance(1e30)
    el = time.perf_counter() - t0
    steps = s2.step
    zu = prob.nx * prob.ny * steps / el
    return {"value": zu, "unit": "zone-updates/s", "cores": threads, "kind": "port",
            "sample": f"KH {prob.nx}x{prob.ny} FP64 RK3, {steps} steps of the oracle restatement of the "
                      f"reference OpenMP path (-O3 -ffast-math -funroll-loops -fopenmp), {el:.1f} s, "
                      f"setup excluded"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seg-rows", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU (--gpus 1): runs across several GPUs are not part of this build",
              file=sys.stderr)
        sys.exit(2)

    from gaukuk_amd import hip
    from gaukuk_amd import problems as P

    log("building KH 2048^2 initial conditions")
    prob = P.kelvin_helmholtz(NX, seed=0, integrator=3, stepmax=-1)
    s = hip.Solver(prob, seg_rows=args.seg_rows)
    log(f"library {hip.version()}, tile {s.tile()}")
    log(f"warmup {args.warmup} steps")
    s.run_steps(args.warmup)
    s.synchronize()

    log(f"timed {args.steps} steps")
    s.synchronize()
    t0 = time.perf_counter()
    total_ms, kms, cnt = s.profile(args.steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    assert s.check_finite()
    zones = NX * NY
    value = zones * args.steps / wall
    kernels = []
    for k, (name, bpc) in enumerate(STAGES):
        avg_ms = kms[k] / max(cnt[k], 1)
        kernels.append({"kernel": name, "launches": cnt[k], "avg_launch_us": avg_ms * 1e3,
                        "bytes_per_launch": bpc * zones,
                        "achieved_GBs": bpc * zones / (avg_ms * 1e-3) / 1e9, "total_ms": kms[k]})
    log(f"wall {wall * 1e3:.1f} ms, events {total_ms:.1f} ms; " +
        ", ".join(f"{q['kernel'].split('<')[1][:28]} {q['avg_launch_us']:.1f} us" for q in kernels) +
        f"; clock kernels {kms[3] / max(cnt[3], 1) * 1e3:.1f} us")
    dom = max(kernels, key=lambda q: q["total_ms"])

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_stage_kernels_kh2048_fp64_rk3.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch", {}).get(dom["kernel"])

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "Kelvin-Helmholtz 2048x2048 FP64, PLM+HLLC+RK3, CFL 0.4, periodic, seeded ICs",
                   "nx": NX, "ny": NY, "integrator": "rk3", "precision": 64},
        "roofline": {"bound": "hbm", "achieved": dom["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["achieved_GBs"] / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": dom["kernel"], "bytes_per_launch": dom["bytes_per_launch"],
                     "avg_launch_us": dom["avg_launch_us"]},
        "stage_kernels": kernels,
        "step_roofline_frac": (320.0 * zones / (wall / args.steps)) / 1e9 / HBM_PEAK_GBS,
    }
    if not args.no_cpu:
        log("CPU baseline (oracle, fast build)")
        out["cpu_baseline"] = cpu_baseline(prob)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
