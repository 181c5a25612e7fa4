"""What the trained networks gain over the method without it, on seeds they were not trained on
(training used seeds 0..63).

gaukuk_amd/data/kh256_flux_mlp_es.npz (the early network fine-tuned a posteriori by evolution
strategies on whole-run errors at t = 0.025, 0.05, 0.1, seeds 0..3, 2580 iterations;
gaukuk_amd.learned.finetune):
* Over a run on the held-out seeds 1000-1002 the gain lasts: 33-35 % (mass), 37-38 % (m_x),
  35-36 % (m_y), 44 % (E) at t = 0.025; 28-32, 32-33, 40-42, 43-45 % at t = 0.05; 8-10, 11-19,
  16-20, 22-24 % at t = 0.1.

gaukuk_amd/data/kh256_flux_mlp_early.npz (trained on frames of t in [0, 0.1), kappa 1):
* Over a run: at t = 0.025 (39 coarse RK3 steps) the corrected FP32 256^2 solver's relative L1
  against the restricted 1024^2 FP64 run is 16-17 % lower than the plain solver's in mass and
  28-31 % lower in m_x, m_y and E on the held-out seeds 1000-1002; at t = 0.05 (78 steps) m_x,
  m_y, E stay 22-31 % lower (mass 3-8 %); by t = 0.1 (155 steps) the gain is gone (the
  instability's chaotic growth then sets the error).

gaukuk_amd/data/kh256_flux_mlp.npz (trained on frames of t in [0, 0.6), kappa 0.5):

* Local: one coarse step from the restricted 1024^2 FP64 state.  The corrected FP32 step's error
  against the restricted fine run is well below the plain FP32 step's (measured 20-35 % lower per
  variable at t0 = 0.15..0.45 on two held-out seeds).
* Over a run (0.3 time units, about 470 coarse steps): the corrected solver stays within a few
  percent of the plain solver's L1 against the restricted 1024^2 run (measured -3 % .. +3 %):
  the local gain does not accumulate into a run-level gain on this problem.
"""
import os

import numpy as np
import pytest

from gaukuk_amd import learned

pytestmark = pytest.mark.gpu

NET = os.path.join(os.path.dirname(learned.__file__), "data", "kh256_flux_mlp.npz")
NET_EARLY = os.path.join(os.path.dirname(learned.__file__), "data", "kh256_flux_mlp_early.npz")
NET_ES = os.path.join(os.path.dirname(learned.__file__), "data", "kh256_flux_mlp_es.npz")


@pytest.fixture(scope="module")
def net():
    from gaukuk_amd import hip
    n = hip.Net.load(NET)
    assert n.kappa == pytest.approx(0.5)
    return n


def test_trained_network_is_a_trained_network(net):
    with np.load(NET) as z:
        assert int(z["seeds"]) == 64 and int(z["n_fine"]) == 1024 and int(z["n_coarse"]) == 256
        loss = z["loss"]
    assert loss[-1, 2] < 0.5 * loss[0, 2]          # training reduced the normalised loss


@pytest.mark.parametrize("t0", [0.15, 0.3])
def test_one_step_error_reduced(net, t0):
    e = learned.one_step_error(net, seed=1001, t0=t0)
    plain, corr = np.array(e["plain"]), np.array(e["corrected"])
    assert np.all(corr < 0.95 * plain), (plain, corr)
    assert np.mean(corr / plain) < 0.85, (plain, corr)


def test_run_level_error_not_degraded(net):
    r = learned.evaluate(net, seeds=(1002,), times=(0.3,), kappas=(0.5,), log=lambda m: None)["seeds"]["1002"]
    plain = np.array(r["fp32_256"]["l1"][0])
    corr = np.array(r["learned_fp32_k0.5_256"]["l1"][0])
    assert r["learned_fp32_k0.5_256"]["reached"]
    assert np.all(corr <= 1.05 * plain), (plain, corr)
    # the 512^2 FP64 solver is the more accurate one: the reference resolves what 256^2 misses
    assert r["fp64_512"]["l1"][0][0] < 0.9 * r["fp64_256"]["l1"][0][0]


def test_early_network_run_level_gain():
    """Run-level gain of the early-time network on a held-out seed at t = 0.025: every conserved
    variable's L1 against the restricted 1024^2 run at least 10 % below the plain FP32 solver's,
    and 20 % on average (measured 17 / 30 / 30 / 29 % for mass, m_x, m_y, E on seed 1002)."""
    from gaukuk_amd import hip
    net = hip.Net.load(NET_EARLY)
    assert net.kappa == pytest.approx(1.0)
    r = learned.evaluate(net, seeds=(1002,), times=(0.025,), kappas=(1.0,), log=lambda m: None)["seeds"]["1002"]
    plain = np.array(r["fp32_256"]["l1"][0])
    corr = np.array(r["learned_fp32_k1_256"]["l1"][0])
    gain = 1 - corr / plain
    print("early network gain at t = 0.025 (mass, m_x, m_y, E):", gain, flush=True)
    assert r["learned_fp32_k1_256"]["reached"]
    assert np.all(gain >= 0.10), gain
    assert np.mean(gain) >= 0.20, gain


def test_finetuned_network_run_level_gain():
    """The a posteriori fine-tuned network on a held-out seed: every conserved variable's L1 against
    the restricted 1024^2 run below the plain FP32 solver's at t = 0.05 (by at least 15 %, 30 % on
    average) and at t = 0.1 (by at least 5 %, 12 % on average; measured on seed 1002: 30 / 33 / 40 /
    44 % and 10 / 17 / 19 / 24 % for mass, m_x, m_y, E)."""
    from gaukuk_amd import hip
    net = hip.Net.load(NET_ES)
    with np.load(NET_ES) as z:
        assert int(z["ft_iters"]) > 0 and float(z["ft_history"][-1, 1]) < float(z["ft_history"][0, 1])
    r = learned.evaluate(net, seeds=(1002,), times=(0.05, 0.1), kappas=(net.kappa,), log=lambda m: None)["seeds"]["1002"]
    key = f"learned_fp32_k{net.kappa:g}_256"
    assert r[key]["reached"]
    for k, (lo, mean) in enumerate(((0.15, 0.30), (0.05, 0.12))):
        gain = 1 - np.array(r[key]["l1"][k]) / np.array(r["fp32_256"]["l1"][k])
        print(f"fine-tuned network gain at t = {(0.05, 0.1)[k]} (mass, m_x, m_y, E):", gain, flush=True)
        assert np.all(gain >= lo), gain
        assert np.mean(gain) >= mean, gain
