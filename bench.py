#!/usr/bin/env python3
"""Benchmark of the per-timestep hydro update on one MI355X.

Workload: Kelvin-Helmholtz 2048x2048, FP64, PLM + HLLC + RK3, CFL 0.4, periodic, seeded ICs.
A step is one full RK3 step (3 fused stage kernels + 1 one-thread clock kernel) over the
whole grid with the state resident in HBM.  Before the --warmup steps a second solver of the same
workload steps untimed for --prewarm-s seconds of wall time, so that the GPU's clocks have settled
when the timed solver starts from its initial state.  Prints one JSON line:
  value         whole-job zone-updates/s over exactly --steps timed steps (enqueued back to back,
                one synchronize)
  roofline      the stage kernel with the most total time, its algorithmic bytes per launch
                divided by its average launch duration from HIP events on the library's stream,
                measured over a second pass of --steps steps right after the timed one
  cpu_baseline  the oracle restatement of the reference's OpenMP CPU path (fast flags) timed on
                this host's cores over a bounded sample of the same workload
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "zone-updates/s & %HBM roofline; time-to-L1-error learned/FP32 vs plain FP64"
HBM_PEAK_GBS = 8000.0
NX = NY = 2048
# The three RK3 stage kernels of a step, with the algorithmic HBM bytes per cell of one launch
# (FP64): the first stage reads U^n and writes U1 (80 B); the two combine stages read U_s and
# U^n and write one state (120 B).
STAGES = [("stage_kernel<double, VanLeer, THREE=false, GRAV=false, LAST=false, VAR=4> (RK3 stage 1)", 80.0),
          ("stage_kernel<double, VanLeer, THREE=true, GRAV=false, LAST=false, VAR=4> (RK3 stage 2)", 120.0),
          ("stage_kernel<double, VanLeer, THREE=true, GRAV=false, LAST=true, VAR=4> (RK3 stage 3)", 120.0)]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


_HOST_CPU = None


def host_cpu():
    """The host cores this process may run on: logical CPUs in its affinity mask, the physical
    cores they belong to, and the CPU model (from /proc/cpuinfo).  Read once, before the first CPU
    baseline: the OpenMP runtime binds the calling thread to one core when it starts (proc_bind
    close), which would shrink the mask seen by later calls."""
    global _HOST_CPU
    if _HOST_CPU is None:
        _HOST_CPU = _read_host_cpu()
    return dict(_HOST_CPU)


def _read_host_cpu():
    cpus = sorted(os.sched_getaffinity(0))
    model, cores, cur = "unknown", set(), {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f.read().split("\n") + [""]:
                if not line.strip():
                    if "processor" in cur and int(cur["processor"]) in cpus:
                        cores.add((cur.get("physical id", "0"), cur.get("core id", cur["processor"])))
                        model = cur.get("model name", model)
                    cur = {}
                    continue
                k, _, v = line.partition(":")
                cur[k.strip()] = v.strip()
    except OSError:
        pass
    return {"logical_cpus": len(cpus), "physical_cores": len(cores) or len(cpus), "model": model}


def cpu_baseline(prob, budget_s: float = 12.0, label: str = "KH", precision: int = 64):
    """Time the oracle (fast build) on this host: a bounded sample of the same workload.  Threads:
    one per physical core of this process's affinity mask, at most OMP_NUM_THREADS when the
    environment sets it (the CPU share of the machine)."""
    host = host_cpu()
    threads = int(os.environ.get("GK_CPU_THREADS", "0")) or host["physical_cores"]
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    from oracle import gko  # test infrastructure: the CPU baseline only
    # one probe step sizes the sample to about budget_s seconds of stepping
    s1 = gko.OracleSim(prob.with_(stepmax=1), precision=precision, flavour="fast")
    t0 = time.perf_counter()
    s1.advance(1e30)
    one = time.perf_counter() - t0
    del s1
    n = max(2, min(200, int(budget_s / max(one, 1e-3))))
    s2 = gko.OracleSim(prob.with_(stepmax=n), precision=precision, flavour="fast")
    t0 = time.perf_counter()
    s2.advance(1e30)
    el = time.perf_counter() - t0
    steps = s2.step
    zu = prob.nx * prob.ny * prob.nz * steps / el
    rk = {1: "Euler", 2: "RK2", 3: "RK3"}[prob.integrator]
    return {"value": zu, "unit": "zone-updates/s", "cores": threads, "kind": "port",
            "sample": f"{label} {prob.nx}x{prob.ny} FP{precision} {rk}, {steps} steps of the oracle restatement of the "
                      f"reference OpenMP path (-O3 -ffast-math -funroll-loops -fopenmp), {el:.1f} s, "
                      f"setup excluded",
            "host": {**host, "omp_threads": threads, "omp_proc_bind": os.environ.get("OMP_PROC_BIND")}}


def prewarm(s, seconds: float, chunk: int = 20):
    """Untimed stepping of solver `s` until `seconds` of wall time have passed.  After an idle
    period the GPU first runs at its boost clock, then its power management pulls the clock down
    for some tens of milliseconds before it settles (stage kernels 150-200 us instead of 125-150 us
    at KH 2048^2); stepping a second solver of the same workload right up to the timed solver's
    warm-up keeps the timed steps out of that transient without changing their initial state."""
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < seconds:
        s.run_steps(chunk)
        s.synchronize()
        steps += chunk
    return {"seconds": time.perf_counter() - t0, "steps": steps,
            "on": "a second solver of the same workload (the timed solver starts from the initial state)"}


def ensemble_l1(probs, steps: int):
    """Relative L1 of the FP32 ensemble against the same ensemble in FP64 after `steps` steps of
    every problem (each with its own dt, from its own state): per problem the largest of the four
    varying conserved variables (rho, m_x, m_y, E); max and median over the problems."""
    from gaukuk_amd import hip
    res = {}
    for prec in (32, 64):
        e = hip.Ensemble(probs, precision=prec)
        e.run_steps(steps)
        e.synchronize()
        res[prec] = e
    worst, per_var = [], [[] for _ in range(4)]
    for q in range(len(probs)):
        a, b = res[32].download(q), res[64].download(q)
        r = [float(np.abs(a[n] - b[n]).sum() / max(np.abs(b[n]).sum(), 1e-300)) for n in (0, 1, 2, 4)]
        worst.append(max(r))
        for k in range(4):
            per_var[k].append(r[k])
    for e in res.values():
        e.close()
    return {"steps": steps, "max": float(np.max(worst)), "median": float(np.median(worst)),
            "median_per_variable": {v: float(np.median(x)) for v, x in zip(("rho", "m_x", "m_y", "E"), per_var)},
            "max_per_variable": {v: float(np.max(x)) for v, x in zip(("rho", "m_x", "m_y", "E"), per_var)}}


def bench_ensemble(args):
    """512 seeded 2D blast problems, 64^2..512^2, FP32, RK2, each with its own dt: aggregate
    zone-updates/s of the batched persistent launches (eager and graph-replayed) against one
    single-grid launch per problem per stage, on the same data."""
    from gaukuk_amd import autotune, hip
    from gaukuk_amd import problems as P
    log("building 512-problem ensemble")
    probs = P.ensemble(512, seed=0)
    ens = hip.Ensemble(probs, precision=32)
    cells = ens.cells
    log(f"{len(probs)} problems, {cells} cells; autotuning")
    best, table = autotune.tune_ensemble(ens, steps=2)
    log(f"tile (seg_rows, waves/CU) = {best}; table {table}")
    for q, p in enumerate(probs):
        ens.upload(q, p.cons)
    ens.run_steps(args.warmup)
    ens.synchronize()
    K = args.steps
    ms_batched = ens.time(K, hip.BATCHED)
    gsteps = 10 if K % 10 == 0 else 1
    ens.capture(gsteps)
    ens.run_steps(gsteps, hip.GRAPH)
    ens.synchronize()
    t0 = time.perf_counter()
    ms_graph = ens.time(K, hip.GRAPH)
    wall_graph = time.perf_counter() - t0
    kl = max(2, K // 10)
    ms_looped = ens.time(kl, hip.LOOPED)
    zu = lambda ms, n: cells * n / (ms * 1e-3)
    per_step_bytes = 100.0 * cells  # FP32 RK2: stage 1 reads + writes 20 B each, stage 2 reads 40 B, writes 20 B
    ens.close()
    log("FP32 vs FP64 relative L1 per problem")
    l1 = ensemble_l1(probs, steps=args.l1_steps)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_ens_stage_kernel_ensemble512_fp32_rk2.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_step")
    out = {
        "metric": METRIC, "value": zu(ms_graph, K), "unit": "zone-updates/s", "n_gpus": 1,
        "steps": K, "warmup": args.warmup, "ms_per_step": ms_graph / K, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "ensemble of 512 seeded 2D blast problems, 64..512 per side, FP32, "
                               "PLM+HLLC+RK2, CFL 0.4, periodic or outflow, one persistent launch per stage, "
                               "graph-captured", "problems": len(probs), "cells": cells,
                   "tile": list(best)},
        "batched_eager_zone_updates_per_s": zu(ms_batched, K),
        "graph_zone_updates_per_s": zu(ms_graph, K),
        # every problem as its own single-grid run: one launch per stage with that grid's automatic
        # plan, as gaukuk_amd.hip.Solver would run it
        "looped_single_grid_zone_updates_per_s": zu(ms_looped, kl),
        "speedup_vs_looped": zu(ms_graph, K) / zu(ms_looped, kl),
        "step_hbm_frac": per_step_bytes / (ms_graph / K * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "wall_s_graph": wall_graph,
        "fp32_vs_fp64_rel_l1": l1,
        # the persistent stage launches dominate the step; their bytes are the whole step's
        "roofline": {"bound": "hbm", "achieved": per_step_bytes / (ms_graph / K * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": per_step_bytes / (ms_graph / K * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": traffic, "kernel": "ens_stage_kernel<float, VanLeer, ...> (both RK2 stages, per step)",
                     "bytes_per_launch": per_step_bytes / 2, "avg_launch_us": ms_graph / K * 1e3 / 2},
    }
    if not args.no_cpu:
        # the oracle on a sample of the ensemble: the largest problems, one after another
        log("CPU baseline (oracle, fast build)")
        big = sorted(probs, key=lambda p: p.nx * p.ny)[-4:]
        done, el = 0, 0.0
        for p in big:
            cb = cpu_baseline(p, budget_s=3.0, label="blast", precision=32)
            done += 1
            el += p.nx * p.ny / cb["value"]
        cb["value"] = sum(p.nx * p.ny for p in big) / el
        cb["sample"] = (f"the 4 largest problems of the ensemble, FP32 RK2, about 3 s each of the oracle "
                        f"restatement of the reference OpenMP path (-O3 -ffast-math -funroll-loops -fopenmp)")
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def bench_rt(args):
    """Rayleigh-Taylor 1024x3072 with gravity, RK2, to t = 0.2 in FP64 and in FP32 (FP64 dt and
    conservation sums in both): zone-updates/s of each and the relative L1 of FP32 vs FP64."""
    from gaukuk_amd import hip
    from gaukuk_amd import problems as P
    prob = P.rayleigh_taylor(1024, 3072, tmax=0.2, dtoutput=0.2)
    res = {}
    for prec in (64, 32):
        s = hip.Solver(prob, precision=prec)
        sums0 = s.sums()
        warm = hip.Solver(prob.with_(tmax=1e30, dtoutput=1e30), precision=prec)
        pw = prewarm(warm, args.prewarm_s / 2)
        t0 = time.perf_counter()
        s.run_to_end()
        s.synchronize()
        wall = time.perf_counter() - t0
        warm.close()
        res[prec] = dict(state=s.download(), steps=s.step, t=s.t, wall=wall, sums0=sums0, sums1=s.sums(), prewarm=pw)
        log(f"RT FP{prec}: {s.step} steps to t={s.t:.6f} in {wall:.3f} s (after {pw['steps']} prewarm steps)")
    a, b = res[32]["state"], res[64]["state"]
    l1 = [float(np.abs(a[n] - b[n]).sum() / max(np.abs(b[n]).sum(), 1e-300)) for n in range(5)]
    zu = {p: prob.nx * prob.ny * res[p]["steps"] / res[p]["wall"] for p in res}
    # per-kernel HIP-event timing of the FP32 stage kernels (RK2 with gravity; a state is 20 B per cell
    # in FP32: stage 1 reads U^n and writes U1 = 40 B/cell, stage 2 reads U1 and U^n and writes = 60 B/cell)
    s = hip.Solver(prob.with_(tmax=1e30, dtoutput=1e30), precision=32)
    prewarm(s, args.prewarm_s / 2)
    _, kms, cnt = s.profile(args.steps)
    cells = prob.nx * prob.ny
    kernels = []
    for k, (name, bpc) in enumerate((("stage_kernel<float, VanLeer, THREE=false, GRAV=true, LAST=false, VAR=0> "
                                       "(RK2 stage 1)", 40.0),
                                      ("stage_kernel<float, VanLeer, THREE=true, GRAV=true, LAST=true, VAR=0> "
                                       "(RK2 stage 2)", 60.0))):
        avg_ms = kms[k] / max(cnt[k], 1)
        kernels.append({"kernel": name, "launches": cnt[k], "avg_launch_us": avg_ms * 1e3,
                        "bytes_per_launch": bpc * cells, "achieved_GBs": bpc * cells / (avg_ms * 1e-3) / 1e9})
    dom = max(kernels, key=lambda q: q["avg_launch_us"] * q["launches"])
    log("FP32 stage kernels: " + ", ".join(f"{q['avg_launch_us']:.1f} us" for q in kernels))
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_stage_kernels_rt1024x3072_rk2.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch", {}).get(dom["kernel"])
    out = {
        "metric": METRIC, "value": zu[64], "unit": "zone-updates/s", "n_gpus": 1,
        "steps": res[64]["steps"], "warmup": args.warmup, "ms_per_step": res[64]["wall"] * 1e3 / res[64]["steps"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "Rayleigh-Taylor 1024x3072, gravity -0.1, periodic x / reflecting y, PLM+HLLC+RK2, "
                               "CFL 0.4, to t = 0.2", "nx": 1024, "ny": 3072},
        "fp32_zone_updates_per_s": zu[32], "fp32_steps": res[32]["steps"],
        "fp32_vs_fp64_rel_l1": l1,
        "mass_drift_fp64": float(res[64]["sums1"][0] / res[64]["sums0"][0] - 1),
        "mass_drift_fp32": float(res[32]["sums1"][0] / res[32]["sums0"][0] - 1),
        "roofline": {"bound": "hbm", "achieved": dom["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["achieved_GBs"] / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom["kernel"],
                     "bytes_per_launch": dom["bytes_per_launch"], "avg_launch_us": dom["avg_launch_us"]},
        "fp32_stage_kernels": kernels,
    }
    if not args.no_cpu:
        log("CPU baseline (oracle, fast build)")
        out["cpu_baseline"] = cpu_baseline(prob, label="Rayleigh-Taylor", precision=64)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def bench_3d(args):
    """The 3D path: a periodic 3D shear layer at 256^3, FP64, PLM+HLLC+RK3 (no BASELINE metric;
    zone-updates/s and the stage kernels' HBM fraction for the design notes)."""
    from gaukuk_amd import hip
    from gaukuk_amd import problems as P
    n = 256
    prob = P.shear3d(n, n, seed=0, stepmax=-1)
    s = hip.Solver(prob, seg_rows=args.seg_rows)
    log(f"3D shear {n}^3, library {hip.version()}, tile {s.tile()}")
    s.run_steps(args.warmup)
    s.synchronize()
    t0 = time.perf_counter()
    total_ms, kms, cnt = s.profile(args.steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    assert s.check_finite()
    zones = n ** 3
    kernels = []
    for k, bpc in enumerate((80.0, 120.0, 120.0)):
        avg_ms = kms[k] / max(cnt[k], 1)
        kernels.append({"kernel": f"stage_kernel<double, VanLeer, ..., VAR=2> (RK3 stage {k + 1})",
                        "avg_launch_us": avg_ms * 1e3, "bytes_per_launch": bpc * zones,
                        "achieved_GBs": bpc * zones / (avg_ms * 1e-3) / 1e9})
    dom = max(kernels, key=lambda q: q["avg_launch_us"])
    out = {"metric": "zone-updates/s (3D path)", "value": zones * args.steps / wall, "unit": "zone-updates/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"3D shear layer {n}^3 FP64, PLM+HLLC+RK3, CFL 0.4, periodic"},
           "roofline": {"bound": "hbm", "achieved": dom["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": dom["achieved_GBs"] / HBM_PEAK_GBS, "traffic": None, "kernel": dom["kernel"],
                        "bytes_per_launch": dom["bytes_per_launch"], "avg_launch_us": dom["avg_launch_us"]},
           "stage_kernels": kernels, "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def time_to_l1(r, key, n: int = 256):
    """Wall-time ratio at equal accuracy: for each output time and conserved variable (mass, m_x,
    m_y, E), the resolution N at which plain FP64 would reach the corrected solver's L1 (at the
    convergence order plain FP64 shows between n^2 and (2n)^2), and the wall time of plain FP64 at
    N (cost ~ N^3: cells x steps, from its measured n^2 run) over the corrected solver's."""
    out = []
    for k in range(len(r[key]["l1"])):
        row = []
        for q in range(4):
            e1, e2, el = r[f"fp64_{n}"]["l1"][k][q], r[f"fp64_{2 * n}"]["l1"][k][q], r[key]["l1"][k][q]
            if not (e1 > e2 > 0 and el > 0):
                row.append(None)
                continue
            order = np.log2(e1 / e2)
            n_eq = n * (e1 / el) ** (1.0 / order)
            row.append({"order": float(order), "fp64_n_equivalent": float(n_eq),
                        "wall_ratio": float((n_eq / n) ** 3 * r[f"fp64_{n}"]["wall_s"] / r[key]["wall_s"])})
        out.append(row)
    return out


def bench_learned(args):
    """Learned flux correction on coarse KH: the corrected FP32 256^2 solver (the early-frame network
    fine-tuned a posteriori on whole-run errors, gaukuk_amd/data/kh256_flux_mlp_es.npz, kappa 1)
    against plain FP64 at 256^2 and 512^2 and plain FP32 at 256^2, from the held-out seed 1000 to
    t = 0.025, 0.05 and 0.1: wall time, steps, relative L1 against the restricted 1024^2 FP64 run,
    and the time-to-equal-L1 estimate.  value = zone-updates/s of the corrected solver; roofline =
    the corrected stage kernel's MLP flops on MFMA."""
    from gaukuk_amd import hip, learned
    net = hip.Net.load(os.path.join(ROOT, "gaukuk_amd", "data", "kh256_flux_mlp_es.npz"))
    n = 256
    prob = learned.kh(1024, 1000)
    coarse = prob.with_(nx=n, ny=n, cons=learned.restrict(prob.cons, 1024 // n))
    # per-launch time of the corrected stage kernel (HIP events on the library's stream)
    s = hip.Solver(coarse, precision=32)
    s.set_net(net)
    prewarm(s, args.prewarm_s / 2)
    K = args.steps
    t0 = time.perf_counter()
    total_ms, kms, cnt = s.profile(K)
    s.synchronize()
    wall = time.perf_counter() - t0
    s.close()
    faces = 2 * n * n                                        # each grid face once
    flops = 2.0 * (16 * 64 + 64 * 64 + 64 * 4) * faces      # MLP per launch (MFMA)
    avg_ms = [kms[k] / max(cnt[k], 1) for k in range(3)]
    dom = int(np.argmax([kms[k] for k in range(3)]))
    log(f"corrected stage kernel per launch: {[round(a * 1e3, 1) for a in avg_ms]} us; "
        f"clock kernels {kms[3] / max(cnt[3], 1) * 1e3:.1f} us")
    times = (0.025, 0.05, 0.1)
    log("accuracy and time to t = 0.025, 0.05, 0.1 against the 1024^2 FP64 run")
    kappa = float(net.kappa)
    ev = learned.evaluate(net, seeds=(1000,), times=times, kappas=(kappa,), log=log)
    r = ev["seeds"]["1000"]
    key = f"learned_fp32_k{kappa:g}_{n}"
    gain = {f"t{t}": [1 - c / p for c, p in zip(r[key]["l1"][k], r["fp32_256"]["l1"][k])]
            for k, t in enumerate(times)}
    ttl = time_to_l1(r, key, n)
    out = {
        "metric": METRIC, "value": n * n * K / wall, "unit": "zone-updates/s", "n_gpus": 1,
        "steps": K, "warmup": args.warmup, "ms_per_step": wall * 1e3 / K, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+bf16", "data": "synthetic",
        "config": {"workload": "Kelvin-Helmholtz 256x256 FP32 with the learned flux correction (MLP 16-64-64-4, "
                               "BF16 MFMA), PLM+HLLC+RK3, CFL 0.4, periodic, seeded noise on 64^2 cells, seed 1000",
                   "nx": n, "ny": n, "kappa": kappa, "network": "kh256_flux_mlp_es.npz"},
        "roofline": {"bound": "mfma", "achieved": flops / (avg_ms[dom] * 1e-3) / 1e12, "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": flops / (avg_ms[dom] * 1e-3) / 1e12 / 2500.0, "traffic": None,
                     "kernel": f"learned_stage_kernel (RK3 stage {dom + 1})", "flops_per_launch": flops,
                     "avg_launch_us": avg_ms[dom] * 1e3},
        "time_to_t": {k: {"wall_s": v["wall_s"], "steps": v["steps"],
                          **{f"l1_t{t}": v["l1"][q] for q, t in enumerate(times)}}
                      for k, v in r.items() if k != "fine"},
        "fine_reference": r["fine"],
        "gain_vs_plain_fp32_256": gain,
        "time_to_equal_l1_fp64_over_learned": {f"t{t}": ttl[q] for q, t in enumerate(times)},
        "wall_ratio_fp64_512_over_learned": r["fp64_512"]["wall_s"] / r[key]["wall_s"],
        "wall_ratio_fp64_256_over_learned": r["fp64_256"]["wall_s"] / r[key]["wall_s"],
        "fallback_faces": r[key].get("faces"),
    }
    if not args.no_cpu:
        log("CPU baseline (oracle, fast build) on the coarse grid")
        out["cpu_baseline"] = cpu_baseline(coarse.with_(stepmax=-1))
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seg-rows", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--l1-steps", type=int, default=20, help="ensemble: steps of the FP32-vs-FP64 comparison")
    ap.add_argument("--prewarm-s", type=float, default=2.0,
                    help="untimed stepping (wall seconds) of a second solver before the warm-up steps")
    ap.add_argument("--workload", default="kh", choices=["kh", "ensemble", "rt", "learned", "3d"],
                    help="kh: the headline (default); ensemble / rt / learned: the other benchmark configurations")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU (--gpus 1): runs across several GPUs are not part of this build",
              file=sys.stderr)
        sys.exit(2)

    from gaukuk_amd import hip
    from gaukuk_amd import problems as P

    if args.workload == "ensemble":
        return bench_ensemble(args)
    if args.workload == "rt":
        return bench_rt(args)
    if args.workload == "3d":
        return bench_3d(args)
    if args.workload == "learned":
        return bench_learned(args)

    log("building KH 2048^2 initial conditions")
    prob = P.kelvin_helmholtz(NX, seed=0, integrator=3, stepmax=-1)
    s = hip.Solver(prob, seg_rows=args.seg_rows)
    warm = hip.Solver(prob, seg_rows=args.seg_rows)
    log(f"library {hip.version()}, tile {s.tile()}")
    pw = prewarm(warm, args.prewarm_s)
    log(f"prewarm {pw['steps']} steps in {pw['seconds']:.2f} s (untimed) on a second solver")
    log(f"warmup {args.warmup} steps")
    s.run_steps(args.warmup)
    s.synchronize()

    # the timed region: exactly --steps steps enqueued back to back, one synchronize
    log(f"timed {args.steps} steps")
    t0 = time.perf_counter()
    s.run_steps(args.steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    zones = NX * NY
    value = zones * args.steps / wall
    # the same number of steps again with a HIP event after every launch (events created before
    # the first launch), for the per-kernel durations
    total_ms, kms, cnt = s.profile(args.steps)
    assert s.check_finite()
    warm.close()
    kernels = []
    for k, (name, bpc) in enumerate(STAGES):
        avg_ms = kms[k] / max(cnt[k], 1)
        kernels.append({"kernel": name, "launches": cnt[k], "avg_launch_us": avg_ms * 1e3,
                        "bytes_per_launch": bpc * zones,
                        "achieved_GBs": bpc * zones / (avg_ms * 1e-3) / 1e9, "total_ms": kms[k]})
    log(f"timed: {wall * 1e3:.2f} ms ({wall * 1e3 / args.steps:.4f} ms/step); event pass {total_ms:.2f} ms; " +
        ", ".join(f"stage {k + 1} {q['avg_launch_us']:.1f} us" for k, q in enumerate(kernels)) +
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
        "event_pass_ms_per_step": total_ms / args.steps,
        "prewarm": pw,
    }
    if not args.no_cpu:
        log("CPU baseline (oracle, fast build)")
        out["cpu_baseline"] = cpu_baseline(prob)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
