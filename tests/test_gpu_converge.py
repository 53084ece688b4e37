"""GPU tests of runs under the reference's inner-iteration control
(main.cpp:1249-1320): per case, running maxima E1..E5 of the norms, the
break when all five normalised norms are below the tolerance or at the
iteration cap, then the physical-time shift.

Bar: the iteration count of every step equals the one the compiled reference
chose, and the final state is bitwise the reference's (norms at 1e-12
relative, as in the other parity tests).  In a pool, cases that end their
inner loop at different iterations each reproduce their own run, in FP64 and
in mixed precision.
"""
import numpy as np
import pytest

from bcm_cgli_amd.configs import CONVERGE
from helpers import describe_mismatch, golden_converge, sha
from test_gpu_parity import STATE, er_close

pytestmark = pytest.mark.gpu

# Flux choices of the FP64 pools (the same alone and pooled: norm row
# layout): update and face halos pushed to the other copy, or fused in place.
FLUX = {"fp64-push": (1, 1, 2), "fp64-inplace": (1, 1, 1), "mixed": None}


def bitwise(a, b, what):
    if not np.array_equal(a.view(np.int64), b.view(np.int64)):
        raise AssertionError(f"{what}: {describe_mismatch(a, b)}")


def ran_rows(iters, max_inner):
    """Norm-history rows of the iterations that ran."""
    return np.concatenate([np.arange(s * max_inner, s * max_inner + n) for s, n in enumerate(iters)])


@pytest.mark.parametrize("name", ["tiny", "ml_tiny", "cfg1", "ml"])
def test_converge_matches_reference(B, name):
    g = golden_converge(name)
    n_steps, max_inner, tol = CONVERGE[name]
    e = B.Engine(0)
    try:
        e.load_case(B.Case(B.case_params(name)))
        iters, er, E = e.run_converge(n_steps, max_inner, tol)
        assert iters.tolist() == g["iters"]
        for n in STATE:
            assert sha(e.get_state(n)) == g["sha256"][f"final_{n}.bin"], n
        rows = ran_rows(iters, max_inner)
        assert er_close(er[rows], g["er"][rows])
        assert not er[np.setdiff1d(np.arange(len(er)), rows)].any()
        assert np.allclose(E, g["er"][:, :5].max(axis=0), rtol=1e-12, atol=0)
    finally:
        e.close()


@pytest.mark.parametrize("fuse", [2, 1])
def test_converge_carries_maxima_across_calls(B, O, fuse):
    # Two calls that pass E on equal one call over all the steps, and fixed-
    # count runs afterwards (captured and eager) continue from the right copy
    # of the iterate: bitwise the oracle's same sequence.
    name = "tiny"
    g = golden_converge(name)
    n_steps, max_inner, tol = CONVERGE[name]
    case = B.Case(B.case_params(name))
    e = B.Engine(0)
    try:
        e.load_case(case)
        e.set_flux(1, 1, fuse)
        it1, _, E = e.run_converge(5, max_inner, tol)
        it2, _, _ = e.run_converge(n_steps - 5, max_inner, tol, E=E)
        assert it1.tolist() + it2.tolist() == g["iters"]
        for n in STATE:
            assert sha(e.get_state(n)) == g["sha256"][f"final_{n}.bin"], n
        e.run(1, 2)
        e.run(1, 1, use_graph=False)
        orc = O.OracleRun(case, nthreads=8)
        orc.run_converge(n_steps, max_inner, tol)
        orc.run(1, 2)
        orc.run(1, 1)
        for n in STATE:
            bitwise(e.get_state(n), orc.state[n], n)
    finally:
        e.close()


def pool(B, cases, slots, mode):
    e = B.Engine(0)
    e.alloc_pool(sum(c.ncube - 1 for c in cases) + slots - len(cases), slots,
                 "mixed" if mode == "mixed" else "fp64")
    for c in cases:
        e.add_case(c)
    if FLUX[mode]:
        e.set_flux(*FLUX[mode])
    return e


def run_alone(B, name, mode):
    n_steps, max_inner, tol = CONVERGE[name]
    e = pool(B, [B.Case(B.case_params(name))], 1, mode)
    try:
        iters, er, E = e.run_converge(n_steps, max_inner, tol)
        st = {n: e.case_state(0, n) for n in ("U1_", "U1", "U1q")}
    finally:
        e.close()
    return iters[:, 0], er[:, 0], E[0], st


@pytest.mark.parametrize("mode", list(FLUX))
def test_pooled_cases_break_on_their_own(B, mode):
    # tiny and ml_tiny end their inner loops at different iterations in
    # steps 4 and 8: the one that ends first is shifted and set aside while
    # the other iterates on.  Placed in either order, each case reproduces its
    # run alone bitwise.
    names = ["tiny", "ml_tiny"]
    n_steps, max_inner, tol = CONVERGE["tiny"]
    assert CONVERGE["ml_tiny"] == (n_steps, max_inner, tol)
    alone = {n: run_alone(B, n, mode) for n in names}
    if mode != "mixed":
        for n in names:
            assert alone[n][0].tolist() == golden_converge(n)["iters"], n
    assert not np.array_equal(alone["tiny"][0], alone["ml_tiny"][0])
    for order in (names, names[::-1]):
        e = pool(B, [B.Case(B.case_params(n)) for n in order], 3, mode)
        try:
            slots = [0, 1]
            iters, er, E = e.run_converge(n_steps, max_inner, tol)
            for s, n in zip(slots, order):
                it1, er1, E1, st1 = alone[n]
                assert iters[:, s].tolist() == it1.tolist(), (order, n)
                bitwise(er[:, s], er1, f"{n} er")
                bitwise(E[s], E1, f"{n} E")
                for k, a in st1.items():
                    bitwise(e.case_state(s, k), a, f"{order} {n} {k}")
            assert not iters[:, 2].any()
        finally:
            e.close()
