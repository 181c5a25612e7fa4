"""Batched ensembles on the GPU: every problem of an ensemble must evolve exactly as the same
problem run alone (same kernel body, its own dt), whatever the launch strategy -- persistent
batched launches, one launch per problem, or a replayed HIP graph -- and whatever the tile."""
import numpy as np
import pytest

from gaukuk_amd import problems as P

pytestmark = pytest.mark.gpu


def small_ensemble(count=10, seed=3, nmin=24, nmax=150):
    return P.ensemble(count, seed=seed, nmin=nmin, nmax=nmax)


@pytest.fixture(scope="module")
def hip():
    from gaukuk_amd import hip as H
    return H


@pytest.mark.parametrize("precision", [32, 64])
def test_ensemble_equals_single_grid_runs(hip, precision):
    probs = small_ensemble()
    print("ensemble sizes", [(p.nx, p.ny, p.bc[0]) for p in probs], flush=True)
    ens = hip.Ensemble(probs, precision=precision)
    ens.run_steps(15)
    ens.synchronize()
    for q, p in enumerate(probs):
        s = hip.Solver(p.with_(stepmax=15), precision=precision).run_to_end()
        t, dt, step = ens.get_time(q)
        assert step == 15 and s.step == 15
        assert t == s.t, (q, t, s.t)
        assert np.array_equal(ens.download(q), s.download()), q
    # each problem keeps its own time step
    dts = {ens.get_time(q)[1] for q in range(len(probs))}
    assert len(dts) > 1


def test_looped_mode_identical(hip):
    probs = small_ensemble(8, seed=5)
    a = hip.Ensemble(probs, precision=32)
    b = hip.Ensemble(probs, precision=32)
    a.run_steps(12, hip.BATCHED)
    b.run_steps(12, hip.LOOPED)
    for q in range(len(probs)):
        assert np.array_equal(a.download(q), b.download(q))


def test_graph_replays_match_eager(hip):
    """Capture 4 steps, replay 3 times with eager work (downloads) between replays; every replay
    must equal the eager run at the same step."""
    probs = small_ensemble(9, seed=7)
    eager = hip.Ensemble(probs, precision=32)
    g = hip.Ensemble(probs, precision=32)
    g.capture(4)
    for rep in range(3):
        eager.run_steps(4, hip.BATCHED)
        g.run_steps(4, hip.GRAPH)
        g.synchronize()
        for q in range(len(probs)):
            assert np.array_equal(g.download(q), eager.download(q)), (rep, q)
            assert g.get_time(q) == eager.get_time(q)


@pytest.mark.parametrize("seg,waves", [(1, 4), (5, 8), (24, 20)])
def test_ensemble_tile_invariance(hip, seg, waves):
    probs = small_ensemble(7, seed=11)
    a = hip.Ensemble(probs, precision=32)
    b = hip.Ensemble(probs, precision=32, seg_rows=seg, waves_per_cu=waves)
    a.run_steps(10)
    b.run_steps(10)
    for q in range(len(probs)):
        assert np.array_equal(a.download(q), b.download(q))


def test_ensemble_fp64_vs_oracle(hip):
    from oracle import gko
    probs = small_ensemble(4, seed=13, nmin=40, nmax=90)
    ens = hip.Ensemble(probs, precision=64)
    ens.run_steps(20)
    for q, p in enumerate(probs):
        ref = gko.OracleSim(p.with_(stepmax=20)).run_to_end()
        a, b = ens.download(q), ref.cons()
        for n in range(3):
            assert np.abs(a[n] - b[n]).sum() <= 1e-12 * max(np.abs(b[n]).sum(), np.abs(b[0]).sum())


def test_autotuner_picks_a_candidate(hip, tmp_path, monkeypatch):
    from gaukuk_amd import autotune
    monkeypatch.setenv("GK_TUNE_CACHE", str(tmp_path / "tune.json"))
    ens = hip.Ensemble(small_ensemble(6, seed=17), precision=32)
    best, table = autotune.tune_ensemble(ens, steps=1, seg_candidates=(8, 16), wave_candidates=(8, 12))
    assert best in table and ens.tile()[:2] == best
    again, cached = autotune.tune_ensemble(ens)
    assert again == best and cached == {}
