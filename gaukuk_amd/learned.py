"""Learned flux correction for coarse Kelvin-Helmholtz runs: data, training, evaluation.

The pipeline, all on the GPU through the C ABI (include/gaukuk.h):

1. ``collect``: for each seed, a 1024^2 FP64 RK3 KH run (the ICs of ``problems.kelvin_helmholtz``)
   is stopped at ``frames`` times in [0, t_end); at each stop ``gk_net_collect`` restricts it 4x
   onto a 256^2 FP64 grid, records one coarse step and the fine run over the same interval, and
   appends one sample per coarse face (features of the restricted state, target = restricted fine
   flux minus coarse flux).
2. ``fit``: input/output scales from the data set's RMS values, then Adam on minibatches drawn
   from the device data set (``gk_net_train``: BF16 MFMA forward/backward, FP32 master weights).
3. ``evaluate``: a held-out seed; L1 error of each coarse solver against the 1024^2 FP64 run
   restricted to the coarse grid, and each solver's wall time to t_end.

Usage (writes the trained network as .npz):
    python -m gaukuk_amd.learned train --seeds 64 --out gaukuk_amd/data/kh256_flux_mlp.npz
    python -m gaukuk_amd.learned evaluate --net gaukuk_amd/data/kh256_flux_mlp.npz
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import problems

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_NET = os.path.join(HERE, "data", "kh256_flux_mlp.npz")
NET_PARAMS = 5508


def restrict(u: np.ndarray, f: int) -> np.ndarray:
    """Conservative average of f x f blocks, [5, ny, nx] -> [5, ny/f, nx/f]."""
    q, ny, nx = u.shape
    return u.reshape(q, ny // f, f, nx // f, f).mean(axis=(2, 4))


def kh(n: int, seed: int) -> problems.Problem:
    """KH ICs on n^2, RK3, run by time (no step cap)."""
    return problems.kelvin_helmholtz(n, seed=seed, integrator=3).with_(stepmax=-1)


def init_params(seed: int) -> np.ndarray:
    """He-uniform weights (last layer x0.1), zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for out, inp, scale in ((64, 16, 1.0), (64, 64, 1.0), (4, 64, 0.1)):
        lim = np.sqrt(6.0 / inp)
        parts += [rng.uniform(-lim, lim, out * inp) * scale, np.zeros(out)]
    return np.concatenate(parts).astype(np.float32)


def collect(net, seeds, n_fine: int = 1024, n_coarse: int = 256, frames: int = 16, t_end: float = 0.6,
            log=print) -> None:
    from . import hip
    f = n_fine // n_coarse
    nfc = (n_coarse + 1) * n_coarse * 2
    net.reserve(net.size() + len(seeds) * frames * nfc)
    for s in seeds:
        t0 = time.time()
        fp = kh(n_fine, s)
        fine = hip.Solver(fp)
        coarse = hip.Solver(fp.with_(nx=n_coarse, ny=n_coarse, cons=restrict(fp.cons, f)))
        for k in range(frames):
            tk = t_end * k / frames
            gap = tk - fine.t
            if gap > 1e-12:
                fine.advance(gap)
            net.collect(fine, coarse)
        fine.close()
        coarse.close()
        log(f"seed {s}: {frames} frames, data set {net.size()} samples, {time.time() - t0:.1f} s")


def fit(net, iters: int = 20000, lr: float = 1e-3, lr_end: float = 1e-5, blocks: int = 256,
        tiles_per_wave: int = 4, seed: int = 0, chunk: int = 500, log=print) -> list:
    """Scales from the data, then Adam with a cosine-decayed learning rate."""
    ms_x, ms_t, w = net.stats()
    in_scale = 1.0 / np.sqrt(np.maximum(ms_x, 1e-12))
    out_scale = np.sqrt(np.maximum(ms_t, 1e-30))
    net.set_scales(in_scale, out_scale)
    net.set_params(init_params(seed))
    history = []
    done = 0
    while done < iters:
        k = min(chunk, iters - done)
        frac = (done + 0.5 * k) / iters
        rate = lr_end + 0.5 * (lr - lr_end) * (1 + np.cos(np.pi * frac))
        losses = net.train(k, float(rate), blocks, tiles_per_wave, seed + 1)
        done += k
        history.append((done, float(rate), float(np.mean(losses))))
        log(f"iter {done}: lr {rate:.2e} loss {np.mean(losses):.4f}")
    return history


def rel_l1(a: np.ndarray, b: np.ndarray) -> list:
    """Relative L1 per conserved variable (mass, m_x, m_y, E)."""
    return [float(np.sum(np.abs(a[q] - b[q])) / np.sum(np.abs(b[q]))) for q in (0, 1, 2, 4)]


def timed_run(prob, precision: int, t_end: float, net=None, repeats: int = 2, stepmax: int = -1):
    """Run prob to t_end (at most stepmax steps); returns (state, wall seconds of the fastest
    repeat, steps, fallbacks, time reached)."""
    from . import hip
    best = None
    out = None
    for _ in range(repeats):
        s = hip.Solver(prob.with_(stepmax=stepmax), precision=precision)
        if net is not None:
            s.set_net(net)
        s.synchronize()
        t0 = time.perf_counter()
        s.advance(t_end)
        s.synchronize()
        dt = time.perf_counter() - t0
        if best is None or dt < best:
            best = dt
        out = (s.download(), s.step, s.net_stats() if net is not None else None, s.t)
        s.close()
    return out[0], best, out[1], out[2], out[3]


def evaluate(net, seed: int = 1000, t_end: float = 0.5, n_fine: int = 1024, n_coarse: int = 256,
             log=print) -> dict:
    """L1 of each coarse solver against the restricted n_fine FP64 run, and wall times."""
    from . import hip
    fp = kh(n_fine, seed)
    truth, t_truth, st_truth, _, _ = timed_run(fp, 64, t_end, repeats=1)
    res = {"seed": seed, "t_end": t_end, "fine": {"n": n_fine, "wall_s": t_truth, "steps": st_truth}}
    ref = restrict(truth, n_fine // n_coarse)
    cases = [("fp64", n_coarse, 64, None), ("fp64", 2 * n_coarse, 64, None),
             ("fp32", n_coarse, 32, None), ("learned_fp32", n_coarse, 32, net)]
    plain_steps = None
    for name, n, prec, nn in cases:
        prob = fp.with_(nx=n, ny=n, cons=restrict(fp.cons, n_fine // n))
        # a corrected run that loses stability would shrink dt without end: cap it
        cap = -1 if nn is None or plain_steps is None else 3 * plain_steps
        u, wall, steps, stats, t = timed_run(prob, prec, t_end, nn, stepmax=cap)
        if name == "fp64" and n == n_coarse:
            plain_steps = steps
        l1 = rel_l1(restrict(u, n // n_coarse), ref)
        key = f"{name}_{n}"
        res[key] = {"n": n, "precision": prec, "wall_s": wall, "steps": steps, "l1": l1,
                    "reached_t_end": bool(abs(t - t_end) < 1e-9)}
        if stats is not None:
            res[key]["faces"] = {"corrected": stats[0], "sensor": stats[1], "positivity": stats[2]}
        log(f"{key}: L1 {np.round(l1, 6).tolist()} wall {wall * 1e3:.1f} ms steps {steps} t {t:.4f}")
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train")
    tr.add_argument("--seeds", type=int, default=64)
    tr.add_argument("--frames", type=int, default=16)
    tr.add_argument("--t-end", type=float, default=0.6)
    tr.add_argument("--n-fine", type=int, default=1024)
    tr.add_argument("--n-coarse", type=int, default=256)
    tr.add_argument("--iters", type=int, default=20000)
    tr.add_argument("--lr", type=float, default=1e-3)
    tr.add_argument("--out", default=DEFAULT_NET)
    tr.add_argument("--eval-t", type=float, default=0.5)
    tr.add_argument("--report", default=None)
    ev = sub.add_parser("evaluate")
    ev.add_argument("--net", default=DEFAULT_NET)
    ev.add_argument("--seed", type=int, default=1000)
    ev.add_argument("--t-end", type=float, default=0.5)
    ev.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    from . import hip
    log = lambda m: print(m, flush=True)
    if a.cmd == "train":
        net = hip.Net()
        t0 = time.time()
        collect(net, list(range(a.seeds)), a.n_fine, a.n_coarse, a.frames, a.t_end, log)
        t1 = time.time()
        hist = fit(net, a.iters, a.lr, log=log)
        t2 = time.time()
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        net.save(a.out, seeds=a.seeds, frames=a.frames, n_fine=a.n_fine, n_coarse=a.n_coarse,
                 iters=a.iters, samples=net.size(), loss=np.array(hist))
        log(f"collect {t1 - t0:.1f} s, train {t2 - t1:.1f} s, saved {a.out}")
        rep = {"collect_s": t1 - t0, "train_s": t2 - t1, "samples": net.size(), "loss": hist}
        rep["eval"] = evaluate(net, t_end=a.eval_t, n_fine=a.n_fine, n_coarse=a.n_coarse, log=log)
    else:
        net = hip.Net.load(a.net)
        rep = evaluate(net, a.seed, a.t_end, log=log)
    if a.report:
        with open(a.report, "w") as fh:
            json.dump(rep, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
