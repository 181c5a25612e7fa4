"""What the trained network (gaukuk_amd/data/kh256_flux_mlp.npz) gains over the method without it,
on a seed it was not trained on (training used seeds 0..63).

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
