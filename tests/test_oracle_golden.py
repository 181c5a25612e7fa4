"""Pin the CPU oracle: bitwise equality with golden vectors made by the compiled reference.

The fixtures come from oracle/_ref (reference sources + a from-file setup plug-in, strict IEEE
flags); see tests/golden/make_golden.py.  The oracle's strict build must reproduce them bit for
bit: final state, final time and step count of whole runs, and every per-function vector.
"""
import os

import numpy as np
import pytest

from oracle import gko
from tests._fixtures import load_run, load_vec, run_cases


@pytest.mark.parametrize("name", run_cases())
def test_whole_run_bitwise(name):
    prob, final, t, steps = load_run(name)
    sim = gko.OracleSim(prob).run_to_end()
    assert sim.step == steps
    assert sim.t == t
    got = sim.cons()
    assert got.shape == final.shape
    assert np.array_equal(got.view(np.uint64), final.view(np.uint64)), \
        f"max abs diff {np.max(np.abs(got - final))}"


@pytest.mark.parametrize("direction", [1, 2, 3])
def test_hllc_vectors(direction):
    v = load_vec("hllc")
    f = gko.hllc(v["hllc_ul"], v["hllc_ur"], direction, 1.4)
    assert np.array_equal(f.view(np.uint64), v[f"hllc_dir{direction}"].view(np.uint64))


def test_plm_vectors():
    v = load_vec("plm")
    ul, ur = gko.plm_row(v["prim"])
    assert np.array_equal(ul.view(np.uint64), v["ul"].view(np.uint64))
    assert np.array_equal(ur.view(np.uint64), v["ur"].view(np.uint64))


def test_eos_vectors():
    v = load_vec("eos")
    assert np.array_equal(gko.cons_to_prim(v["cons"], 1.4).view(np.uint64), v["c2p"].view(np.uint64))
    assert np.array_equal(gko.prim_to_cons(v["p2c_in"], 1.4).view(np.uint64), v["p2c"].view(np.uint64))
    assert gko.cmax(v["cons"], 1.4) == float(v["cmax"][0])


def test_random_configurations_match_reference():
    """1000 seeded random configurations (1D/2D/3D, every face type, integrator, order, gravity,
    point masses, inner region, immersed shell): the oracle's step count, final time and state
    checksums equal the compiled reference's."""
    from tests._random_cases import CASES, checksums, describe, random_problem
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "random_configs.npz"))
    assert len(ref["steps"]) == CASES
    for seed in range(CASES):
        prob = random_problem(seed)
        sim = gko.OracleSim(prob).run_to_end()
        assert sim.step == ref["steps"][seed], describe(seed, prob)
        assert sim.t == ref["t"][seed], describe(seed, prob)
        np.testing.assert_allclose(checksums(sim.cons()), ref["sums"][seed], rtol=1e-14, atol=1e-300,
                                   err_msg=describe(seed, prob))
