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
4. ``finetune``: a posteriori, on the error of whole corrected runs rather than of one step:
   evolution strategies on the run-level objective (below), the corrected GPU solver being the
   only model evaluated (no adjoint of the stage is needed).

Usage (writes the trained network as .npz):
    python -m gaukuk_amd.learned train --seeds 64 --out gaukuk_amd/data/kh256_flux_mlp.npz
    python -m gaukuk_amd.learned evaluate --net gaukuk_amd/data/kh256_flux_mlp.npz
    python -m gaukuk_amd.learned finetune --net gaukuk_amd/data/kh256_flux_mlp_early.npz \
        --out gaukuk_amd/data/kh256_flux_mlp_es.npz
A fine-tuned network given as --net continues its fine-tuning (the saved record adds up the
iterations and joins the objective histories); the shipped kh256_flux_mlp_es.npz is 120 + 260 +
300 + 400 + 600 iterations from the early network at the defaults, then 400 + 500 with
--sigma 0.01 --lr 0.005.
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


# The seeded perturbation is drawn on a 64^2 grid and held constant over its cells, so every grid
# of 64^2 or finer starts from the same state and the fine run is a converged reference for the
# coarse ones (noise drawn per fine cell cannot be represented by a coarse grid, and its chaotic
# growth then dominates any coarse-vs-fine error).
NOISE_CELLS = 64


def kh(n: int, seed: int) -> problems.Problem:
    """KH ICs on n^2, RK3, run by time (no step cap); seeded noise on the 64^2 grid."""
    return problems.kelvin_helmholtz(n, seed=seed, integrator=3, noise_cells=NOISE_CELLS).with_(stepmax=-1)


def init_params(seed: int) -> np.ndarray:
    """He-uniform weights (last layer x0.1), zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for out, inp, scale in ((64, 16, 1.0), (64, 64, 1.0), (4, 64, 0.1)):
        lim = np.sqrt(6.0 / inp)
        parts += [rng.uniform(-lim, lim, out * inp) * scale, np.zeros(out)]
    return np.concatenate(parts).astype(np.float32)


def collect(net, seeds, n_fine: int = 1024, n_coarse: int = 256, frames: int = 16, t_end: float = 0.6,
            log=print, window: int = 1) -> None:
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
            net.collect(fine, coarse, window)
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


def timed_run(prob, precision: int, times, net=None, repeats: int = 2, stepmax: int = -1):
    """Run prob through the output times (ascending; at most stepmax steps).  Returns the states
    at those times, the wall seconds of advancing to the last one (fastest repeat, downloads
    excluded), the step count, the fallback counts and the time reached."""
    from . import hip
    best = None
    out = None
    for _ in range(repeats):
        s = hip.Solver(prob.with_(stepmax=stepmax), precision=precision)
        if net is not None:
            s.set_net(net)
        s.synchronize()
        wall = 0.0
        states = []
        for t in times:
            t0 = time.perf_counter()
            if s.time()[3]:
                s.advance(t - s.t)
            s.synchronize()
            wall += time.perf_counter() - t0
            states.append(s.download())
        if best is None or wall < best:
            best = wall
        out = (states, s.step, s.net_stats() if net is not None else None, s.t)
        s.close()
    return out[0], best, out[1], out[2], out[3]


def evaluate(net, seeds=(1000,), times=(0.3, 0.5), kappas=(0.5,), n_fine: int = 1024, n_coarse: int = 256,
             log=print) -> dict:
    """For held-out seeds: relative L1 (mass, m_x, m_y, E) at each output time of the plain FP64
    solver at n_coarse and 2 n_coarse, the plain FP32 solver and the corrected FP32 solver (one per
    kappa) at n_coarse, all restricted to n_coarse, against the n_fine FP64 run; and wall times."""
    res = {"n_fine": n_fine, "n_coarse": n_coarse, "times": list(times), "seeds": {}}
    for seed in seeds:
        fp = kh(n_fine, seed)
        truth, t_truth, st_truth, _, _ = timed_run(fp, 64, times, repeats=1)
        refs = [restrict(u, n_fine // n_coarse) for u in truth]
        r = {"fine": {"n": n_fine, "wall_s": t_truth, "steps": st_truth}}
        cases = [("fp64", n_coarse, 64, None), ("fp64", 2 * n_coarse, 64, None), ("fp32", n_coarse, 32, None)]
        cases += [(f"learned_fp32_k{k:g}", n_coarse, 32, k) for k in kappas]
        plain_steps = None
        for name, n, prec, kappa in cases:
            prob = fp.with_(nx=n, ny=n, cons=restrict(fp.cons, n_fine // n))
            if kappa is not None:
                net.kappa = kappa
            # a corrected run that loses stability would shrink dt without end: cap it
            cap = -1 if kappa is None or plain_steps is None else 3 * plain_steps
            us, wall, steps, stats, t = timed_run(prob, prec, times, net if kappa is not None else None, stepmax=cap)
            if name == "fp64" and n == n_coarse:
                plain_steps = steps
            l1 = [rel_l1(restrict(u, n // n_coarse), ref) for u, ref in zip(us, refs)]
            key = f"{name}_{n}"
            r[key] = {"n": n, "precision": prec, "wall_s": wall, "steps": steps, "l1": l1,
                      "reached": bool(abs(t - times[-1]) < 1e-9)}
            if stats is not None:
                r[key]["kappa"] = kappa
                r[key]["faces"] = {"corrected": stats[0], "sensor": stats[1], "positivity": stats[2]}
            log(f"seed {seed} {key}: wall {wall * 1e3:.1f} ms steps {steps} " +
                " ".join(f"t={tt}: {np.round(x, 5).tolist()}" for tt, x in zip(times, l1)))
        res["seeds"][str(seed)] = r
    return res


# Parameter blocks of the network (W1, b1, W2, b2, W3, b3): each weight block and its layer's bias
# are perturbed on the scale of that layer's RMS weight.
_BLOCKS = ((0, 1024, 1088), (1088, 5184, 5248), (5248, 5504, 5508))


def finetune(net, seeds=(0, 1, 2, 3), times=(0.025, 0.05, 0.1), iters: int = 120, pairs: int = 12,
             sigma: float = 0.02, lr: float = 0.01, seed: int = 0, n_fine: int = 1024, n_coarse: int = 256,
             log=print) -> list:
    """A posteriori fine-tuning by evolution strategies (antithetic Gaussian perturbations, rank
    utilities, Adam).  Objective: the corrected FP32 solver's relative L1 (mass, m_x, m_y, E) against
    the restricted n_fine FP64 run, divided by the plain FP32 solver's, averaged over variables,
    output times and seeds (1 = no gain).  A run that does not reach the last time (its dt collapsed
    under a step cap of 3x the plain run's steps) scores 10 per variable.  Keeps the best parameters
    seen on the unperturbed objective and leaves them in `net`.  Returns [[iteration, objective]]."""
    from . import hip
    f = n_fine // n_coarse
    refs, probs, plain = {}, {}, {}
    for s in seeds:
        fp = kh(n_fine, s)
        us, _, _, _, _ = timed_run(fp, 64, times, repeats=1)
        refs[s] = [restrict(u, f) for u in us]
        probs[s] = fp.with_(nx=n_coarse, ny=n_coarse, cons=restrict(fp.cons, f))
        pu, _, steps, _, _ = timed_run(probs[s], 32, times, repeats=1)
        plain[s] = [np.array(rel_l1(u, r)) for u, r in zip(pu, refs[s])]
        probs[s] = probs[s].with_(stepmax=3 * steps)

    def objective(theta):
        net.set_params(theta.astype(np.float32))
        tot = 0.0
        for s in seeds:
            sv = hip.Solver(probs[s], precision=32)
            sv.set_net(net)
            for k, t in enumerate(times):
                if sv.time()[3]:
                    sv.advance(t - sv.t)
                e = np.array(rel_l1(sv.download(), refs[s][k])) if abs(sv.t - t) < 1e-9 else np.full(4, 10.0)
                tot += float(np.mean(e / plain[s][k]))
            sv.close()
        return tot / (len(seeds) * len(times))

    theta = net.params().astype(np.float64)
    scale = np.empty_like(theta)
    for a, b, c in _BLOCKS:
        scale[a:c] = np.sqrt(np.mean(theta[a:b] ** 2)) + 1e-8
    rng = np.random.default_rng(seed)
    m, v = np.zeros_like(theta), np.zeros_like(theta)
    best = (objective(theta), theta.copy())
    hist = [[0, best[0]]]
    log(f"finetune: start objective {best[0]:.4f}")
    for it in range(1, iters + 1):
        eps = rng.standard_normal((pairs, theta.size))
        fs = np.array([objective(theta + sigma * scale * e) for e in eps] +
                      [objective(theta - sigma * scale * e) for e in eps])
        ranks = np.empty(2 * pairs)
        ranks[np.argsort(fs)] = np.arange(2 * pairs)
        util = 0.5 - ranks / (2 * pairs - 1)           # the lowest objective gets +0.5
        g = -((util[:pairs] - util[pairs:])[:, None] * eps).sum(0) / pairs
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        theta = theta - lr * sigma * scale * (m / (1 - 0.9 ** it)) / (np.sqrt(v / (1 - 0.999 ** it)) + 1e-8)
        if it % 5 == 0 or it == iters:
            fo = objective(theta)
            hist.append([it, fo])
            if fo < best[0]:
                best = (fo, theta.copy())
            log(f"finetune iter {it}: objective {fo:.4f}, best {best[0]:.4f}")
    net.set_params(best[1].astype(np.float32))
    return hist


def one_step_error(net, seed: int = 1000, t0: float = 0.3, n_fine: int = 1024, n_coarse: int = 256) -> dict:
    """Local (one coarse step) error of the plain and the corrected FP32 coarse solvers: from the
    restricted state of the n_fine FP64 run at t0, one coarse step of length D (both solvers take
    the same D, from the same state) against the restricted fine run over D.  Per variable (mass,
    m_x, m_y, E): L1 of (coarse - restricted fine) over L1 of the true change."""
    from . import hip
    fp = kh(n_fine, seed)
    f = n_fine // n_coarse
    fine = hip.Solver(fp)
    fine.advance(t0)
    uf = fine.download()
    u0 = restrict(uf, f)
    cp = fp.with_(nx=n_coarse, ny=n_coarse, cons=u0)
    plain = hip.Solver(cp, precision=32)
    plain.run_steps(1)
    up, D = plain.download(), plain.time()[1]
    corr = hip.Solver(cp, precision=32)
    corr.set_net(net)
    corr.run_steps(1)
    uc = corr.download()
    fine.advance(D)
    ref = restrict(fine.download(), f)

    def err(u):
        return [float(np.sum(np.abs(u[q] - ref[q])) / np.sum(np.abs(ref[q] - u0[q]))) for q in (0, 1, 2, 4)]

    return {"seed": seed, "t0": t0, "dt": D, "plain": err(up), "corrected": err(uc)}


def continued_record(meta: dict, hist, iters: int) -> dict:
    """The metadata to save after `iters` fine-tuning iterations with objective history `hist`
    ([iteration, objective] pairs) on a network whose saved metadata is `meta`.  A network
    fine-tuned before continues that record: the iterations add up and the histories are joined
    (this run's iterations offset by the earlier total); the earlier run's other ft_* settings are
    dropped, for the caller to replace with this run's."""
    prev_iters = int(meta.get("ft_iters", 0))
    prev_hist = np.asarray(meta.get("ft_history", np.zeros((0, 2))), dtype=np.float64).reshape(-1, 2)
    out = {k: v for k, v in meta.items() if not k.startswith("ft_")}
    new = np.array([[prev_iters + h[0], h[1]] for h in hist], dtype=np.float64).reshape(-1, 2)
    out["ft_iters"] = prev_iters + int(iters)
    out["ft_history"] = np.concatenate([prev_hist, new])
    return out


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
    tr.add_argument("--eval-seeds", default="1000,1001,1002")
    tr.add_argument("--eval-times", default="0.3,0.5")
    tr.add_argument("--kappas", default="0.25,0.5,1.0")
    tr.add_argument("--kappa", type=float, default=0.5)
    tr.add_argument("--window", type=int, default=1, help="coarse steps per target window")
    tr.add_argument("--report", default=None)
    ev = sub.add_parser("evaluate")
    ev.add_argument("--net", default=DEFAULT_NET)
    ev.add_argument("--eval-seeds", default="1000,1001,1002")
    ev.add_argument("--eval-times", default="0.3,0.5")
    ev.add_argument("--kappas", default="0.25,0.5,1.0")
    ev.add_argument("--report", default=None)
    ft = sub.add_parser("finetune")
    ft.add_argument("--net", default=os.path.join(HERE, "data", "kh256_flux_mlp_early.npz"))
    ft.add_argument("--out", default=os.path.join(HERE, "data", "kh256_flux_mlp_es.npz"))
    ft.add_argument("--seeds", default="0,1,2,3")
    ft.add_argument("--times", default="0.025,0.05,0.1")
    ft.add_argument("--iters", type=int, default=120)
    ft.add_argument("--pairs", type=int, default=12)
    ft.add_argument("--sigma", type=float, default=0.02)
    ft.add_argument("--lr", type=float, default=0.01)
    ft.add_argument("--eval-seeds", default="1000,1001,1002")
    ft.add_argument("--eval-times", default="0.025,0.05,0.1")
    ft.add_argument("--kappas", default=None)
    ft.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    from . import hip
    log = lambda m: print(m, flush=True)
    seeds = [int(v) for v in a.eval_seeds.split(",")]
    times = [float(v) for v in a.eval_times.split(",")]
    kappas = [float(v) for v in a.kappas.split(",")] if a.kappas else None
    if a.cmd == "finetune":
        net = hip.Net.load(a.net)
        kappas = kappas or [net.kappa]
        ft_seeds = [int(v) for v in a.seeds.split(",")]
        ft_times = [float(v) for v in a.times.split(",")]
        t0 = time.time()
        hist = finetune(net, ft_seeds, ft_times, a.iters, a.pairs, a.sigma, a.lr, log=log)
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        with np.load(a.net, allow_pickle=False) as z:
            meta = {k: z[k] for k in z.files if k not in ("params", "in_scale", "out_scale", "kappa")}
        meta = continued_record(meta, hist, a.iters)
        net.save(a.out, **meta, ft_seeds=np.array(ft_seeds), ft_times=np.array(ft_times),
                 ft_pairs=a.pairs, ft_sigma=a.sigma, ft_lr=a.lr)
        log(f"finetune {time.time() - t0:.1f} s, saved {a.out}")
        rep = {"finetune_s": time.time() - t0, "history": hist, "eval": evaluate(net, seeds, times, kappas, log=log)}
    elif a.cmd == "train":
        net = hip.Net()
        net.kappa = a.kappa
        t0 = time.time()
        collect(net, list(range(a.seeds)), a.n_fine, a.n_coarse, a.frames, a.t_end, log, a.window)
        t1 = time.time()
        hist = fit(net, a.iters, a.lr, log=log)
        t2 = time.time()
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        net.save(a.out, seeds=a.seeds, frames=a.frames, n_fine=a.n_fine, n_coarse=a.n_coarse,
                 iters=a.iters, samples=net.size(), loss=np.array(hist), noise_cells=NOISE_CELLS,
                 t_end=a.t_end, window=a.window)
        log(f"collect {t1 - t0:.1f} s, train {t2 - t1:.1f} s, saved {a.out}")
        rep = {"collect_s": t1 - t0, "train_s": t2 - t1, "samples": net.size(), "loss": hist}
        rep["eval"] = evaluate(net, seeds, times, kappas, a.n_fine, a.n_coarse, log=log)
        net.kappa = a.kappa
        rep["one_step"] = [one_step_error(net, sd, t0, a.n_fine, a.n_coarse) for sd in seeds for t0 in (0.15, 0.3, 0.45)]
        for e in rep["one_step"]:
            log(f"one step seed {e['seed']} t0={e['t0']}: plain {np.round(e['plain'], 4).tolist()} "
                f"corrected {np.round(e['corrected'], 4).tolist()}")
    else:
        net = hip.Net.load(a.net)
        k0 = net.kappa
        rep = evaluate(net, seeds, times, kappas, log=log)
        net.kappa = k0
        rep["one_step"] = [one_step_error(net, sd, t0) for sd in seeds for t0 in (0.15, 0.3, 0.45)]
        for e in rep["one_step"]:
            log(f"one step seed {e['seed']} t0={e['t0']}: plain {np.round(e['plain'], 4).tolist()} "
                f"corrected {np.round(e['corrected'], 4).tolist()}")
    if a.report:
        with open(a.report, "w") as fh:
            json.dump(rep, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
