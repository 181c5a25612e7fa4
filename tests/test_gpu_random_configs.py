"""Seeded random configurations (tests/_random_cases.py), GPU against the oracle: per-variable
relative L1 (FP64: 1e-12; FP32: 5e-5), step counts and final times.  The oracle reproduces the
compiled reference bit for bit on every case (test_oracle_golden pins its checksums against the
reference's, written by tests/golden/make_golden.py --random).
"""
import numpy as np
import pytest

from tests._random_cases import CASES, describe, random_problem, with_outputs


def rel_l1(a, b):
    scale = np.sum(np.abs(b[0]))
    return np.array([np.sum(np.abs(a[n] - b[n])) / (np.sum(np.abs(b[n])) or scale) for n in range(5)])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("seed", range(CASES))
def test_random_config_vs_oracle(seed, precision):
    from gaukuk_amd import hip
    from oracle import gko
    prob = random_problem(seed)
    print(describe(seed, prob), f"precision={precision}", flush=True)
    ref = gko.OracleSim(prob, precision=precision).run_to_end()
    s = hip.Solver(prob, precision=precision).run_to_end()
    assert s.step == ref.step
    tol = 1e-12 if precision == 64 else 5e-5
    assert abs(s.t - ref.t) <= tol * ref.t
    err = rel_l1(s.download(), ref.cons().astype(np.float64))
    assert np.all(err <= tol), err


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(150))
def test_random_config_outputs_vs_oracle(seed):
    """FP64 runs to a random end time in several outputs.  (FP32 is compared at a step count only:
    the oracle keeps the reference's Real-typed clock, whose loop need not end on time in float.)"""
    from gaukuk_amd import hip
    from oracle import gko
    prob = with_outputs(random_problem(seed), seed)
    print(describe(seed, prob), f"tmax={prob.tmax} dtoutput={prob.dtoutput} stepmax={prob.stepmax}", flush=True)
    ref = gko.OracleSim(prob).run_to_end()
    s = hip.Solver(prob).run_to_end()
    assert s.step == ref.step
    assert abs(s.t - ref.t) <= 1e-12 * ref.t
    err = rel_l1(s.download(), ref.cons())
    assert np.all(err <= 1e-12), err
