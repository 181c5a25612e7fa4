"""Tile autotuner: times candidate launch shapes on the device and keeps the fastest.

Single grids: rows per wave segment of the stage kernel.  Ensembles: rows per segment and the
resident waves per CU of the persistent grid.  The choice never changes results (the kernels are
bitwise tile-invariant), only speed.  Choices are cached in a JSON file keyed by the problem
signature ($GK_TUNE_CACHE, default ~/.cache/gaukuk_amd/tune.json); an unwritable cache is skipped.
"""
from __future__ import annotations

import json
import os

SEG_CANDIDATES = (8, 12, 16, 24, 32)
WAVE_CANDIDATES = (8, 12, 16, 20)


def _cache_path() -> str:
    return os.environ.get("GK_TUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "gaukuk_amd", "tune.json")


def _load() -> dict:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key: str, value) -> None:
    data = _load()
    data[key] = value
    try:
        os.makedirs(os.path.dirname(_cache_path()), exist_ok=True)
        with open(_cache_path(), "w") as f:
            json.dump(data, f, indent=1)
    except OSError:
        pass


def ensemble_key(ens) -> str:
    p0 = ens.probs[0]
    return f"ens:{ens.precision}:{p0.integrator}:{p0.rc_order}:{len(ens)}:{ens.cells}"


def tune_ensemble(ens, steps: int = 2, seg_candidates=SEG_CANDIDATES, wave_candidates=WAVE_CANDIDATES,
                  use_cache: bool = True):
    """Pick (seg_rows, waves_per_cu) for `ens` by timing `steps` batched steps per candidate.
    Advances the ensemble's state; re-upload initial data afterwards if it matters.
    Returns (best, table) with table = {(seg, waves): ms_per_step}."""
    key = ensemble_key(ens)
    if use_cache:
        hit = _load().get(key)
        if hit:
            ens.set_tile(*hit["best"])
            return tuple(hit["best"]), {}
    table = {}
    for s in seg_candidates:
        for w in wave_candidates:
            ens.set_tile(s, w)
            ens.run_steps(1)
            table[(s, w)] = ens.time(steps) / steps
    best = min(table, key=table.get)
    ens.set_tile(*best)
    _store(key, {"best": list(best), "ms_per_step": table[best]})
    return best, table


def tune_solver(solver, steps: int = 5, seg_candidates=(12, 16, 18, 24, 32), use_cache: bool = True):
    """Pick rows per segment for a single-grid Solver by timing `steps` profiled steps each."""
    p = solver.prob
    key = f"grid:{solver.precision}:{p.integrator}:{p.rc_order}:{p.nx}x{p.ny}"
    if use_cache:
        hit = _load().get(key)
        if hit:
            solver._check(solver.L.gk_set_tile(solver.h, hit["best"]))
            return hit["best"], {}
    table = {}
    for s in seg_candidates:
        solver._check(solver.L.gk_set_tile(solver.h, s))
        solver.run_steps(1)
        total_ms, _, _ = solver.profile(steps)
        table[s] = total_ms / steps
    best = min(table, key=table.get)
    solver._check(solver.L.gk_set_tile(solver.h, best))
    _store(key, {"best": best, "ms_per_step": table[best]})
    return best, table
