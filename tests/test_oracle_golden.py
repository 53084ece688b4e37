"""The FP64 CPU oracle against the reference solver compiled from its own sources.

Bitwise: state after each of the first three RK stages and after the whole
run, and the per-inner-iteration norms er[1..7] (the oracle sums in the
reference's sequential order, so even the norms match exactly).
"""
import numpy as np
import pytest

from bcm_cgli_amd.configs import CONVERGE, STEPPING
from helpers import golden, golden_converge, sha

STATE = ("U1_", "U1", "U1q", "U1p1", "U1p2", "Rku1", "Residual1")


def run_oracle(B, O, name, stage_hashes=False):
    case = B.Case(B.case_params(name))
    orc = O.OracleRun(case, nthreads=8)
    n_steps, n_inner = STEPPING[name]
    stages = {}
    if stage_hashes:
        for RK in (1, 2, 3):
            orc.stage(RK)
            for n in ("U1_", "Rku1", "Residual1"):
                stages[f"stage{RK}_{n}.bin"] = sha(orc.state[n])
        # restart from the initial state for the full run
        orc = O.OracleRun(case, nthreads=8)
    hist = orc.run(n_steps, n_inner)
    return orc, hist, stages


def test_oracle_tiny_bitwise(B, O):
    g = golden("tiny")
    orc, hist, stages = run_oracle(B, O, "tiny", stage_hashes=True)
    for k, v in stages.items():
        assert v == g["sha256"][k], k
    for n in STATE:
        assert sha(orc.state[n]) == g["sha256"][f"final_{n}.bin"], n
    assert np.array_equal(hist.view(np.int64), g["er"].view(np.int64))


@pytest.mark.slow
def test_oracle_cfg1_bitwise(B, O):
    g = golden("cfg1")
    orc, hist, stages = run_oracle(B, O, "cfg1", stage_hashes=True)
    for k, v in stages.items():
        assert v == g["sha256"][k], k
    for n in STATE:
        assert sha(orc.state[n]) == g["sha256"][f"final_{n}.bin"], n
    assert np.array_equal(hist.view(np.int64), g["er"].view(np.int64))


@pytest.mark.parametrize("name", ["ml_tiny", "ml", "ml3"])
def test_oracle_two_level_bitwise(B, O, name):
    # Grids of two and three cube sizes (ml3: edges 0.25, 0.125, 0.0625, the
    # finest block inside the middle one): injection into small cubes' ghosts
    # and 8-cell means into big cubes' ghosts across every size change.
    g = golden(name)
    orc, hist, stages = run_oracle(B, O, name, stage_hashes=True)
    for k, v in stages.items():
        assert v == g["sha256"][k], k
    for n in STATE:
        assert sha(orc.state[n]) == g["sha256"][f"final_{n}.bin"], n
    assert np.array_equal(hist.view(np.int64), g["er"].view(np.int64))


@pytest.mark.parametrize("name", ["tiny", "ml_tiny", "cfg1", "ml"])
def test_oracle_convergence_break_bitwise(B, O, name):
    # The reference's inner-iteration control (running-max normalised norms
    # below the tolerance, or the iteration cap) ends each step's inner loop
    # at the iteration the reference ends it, and the runs agree bitwise.
    g = golden_converge(name)
    n_steps, max_inner, tol = CONVERGE[name]
    assert len(set(g["iters"])) > 2 and max(g["iters"]) == max_inner
    orc = O.OracleRun(B.Case(B.case_params(name)), nthreads=8)
    iters, hist, E = orc.run_converge(n_steps, max_inner, tol)
    assert iters.tolist() == g["iters"]
    for n in STATE:
        assert sha(orc.state[n]) == g["sha256"][f"final_{n}.bin"], n
    assert np.array_equal(hist.view(np.int64), g["er"].view(np.int64))
    # E1..E5 are the running maxima of the norms over every iteration run.
    assert np.array_equal(E, g["er"][:, :5].max(axis=0))
