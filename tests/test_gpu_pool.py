"""GPU tests of the case pool: several independent cases resident on one
device and run by the same kernel launches.

Bar: every state array and every norm of a case is bitwise the same whether
the case runs alone or next to other cases, wherever it sits in the pool and
whatever came and went around it; a pooled case placed after another one
still reproduces the reference's own final state.
"""
import numpy as np
import pytest

from bcm_cgli_amd.configs import STEPPING, batch_cases
from helpers import describe_mismatch, golden, sha

pytestmark = pytest.mark.gpu

FLUX = (1, 1, True)   # the same flux choice alone and pooled (norm row layout)


def small_cases():
    # 8- and 27-cube cases with differing M, Re and sphere position.
    return batch_cases(3, seed=77, lattices=(2, 3), ico_level=3)


def bitwise(a, b, what):
    if not np.array_equal(a.view(np.int64), b.view(np.int64)):
        raise AssertionError(f"{what}: {describe_mismatch(a, b)}")


def run_alone(B, params, n_steps, n_inner):
    case = B.Case(params)
    e = B.Engine(0)
    try:
        e.load_case(case)
        e.set_flux(*FLUX)
        er = e.run(n_steps, n_inner)
        st = {n: e.get_state(n) for n in ("U1_", "U1", "U1q", "Residual1")}
    finally:
        e.close()
    return st, er


def test_pool_matches_alone(B):
    params = small_cases()
    n_steps, n_inner = 2, 2
    alone = [run_alone(B, p, n_steps, n_inner) for p in params]
    cases = [B.Case(p) for p in params]
    e = B.Engine(0)
    try:
        e.alloc_pool(sum(c.ncube - 1 for c in cases) + 5, 4)
        slots = [e.add_case(c) for c in cases]
        e.set_flux(*FLUX)
        info = e.pool_info()
        assert info["active"] == 3 and info["used_cubes"] == sum(c.ncube - 1 for c in cases)
        er = e.run(n_steps, n_inner)
        assert er.shape == (n_steps * n_inner, 4, 7)
        for s, (st, er1) in zip(slots, alone):
            for n, a in st.items():
                bitwise(e.case_state(s, n), a, f"slot {s} {n}")
            bitwise(er[:, s, :], er1, f"slot {s} er")
    finally:
        e.close()


def test_pool_admit_and_retire(B):
    pa, pb, pc = small_cases()
    assert pa["lattice"][0] <= pb["lattice"][0]
    ca, cb, cc = B.Case(pa), B.Case(pb), B.Case(pc)
    e = B.Engine(0)
    try:
        e.alloc_pool((ca.ncube - 1) + (cb.ncube - 1) + (cc.ncube - 1), 3)
        sa = e.add_case(ca)
        sb = e.add_case(cb)
        e.set_flux(*FLUX)
        e.run(1, 2)
        e.remove_case(sa)
        assert e.pool_info()["active"] == 1
        sc = e.add_case(cc)   # first fit: lands where it fits, possibly in sa's range
        e.run(1, 2)
        st_b, _ = run_alone(B, pb, 2, 2)
        st_c, _ = run_alone(B, pc, 1, 2)
        for n in ("U1_", "U1", "U1q"):
            bitwise(e.case_state(sb, n), st_b[n], f"B {n}")
            bitwise(e.case_state(sc, n), st_c[n], f"C {n}")
    finally:
        e.close()


def test_pooled_case_matches_reference_fixture(B):
    """The tiny case placed behind another case (pool offset != 0) still
    reproduces the reference's final state."""
    g = golden("tiny")
    other = B.Case(small_cases()[1])
    tiny = B.Case(B.case_params("tiny"))
    e = B.Engine(0)
    try:
        e.alloc_pool((other.ncube - 1) + (tiny.ncube - 1), 2)
        e.add_case(other)
        s = e.add_case(tiny)
        n_steps, n_inner = STEPPING["tiny"]
        e.run(n_steps, n_inner)
        for n in ("U1_", "U1", "U1q", "Residual1"):
            assert sha(e.case_state(s, n)) == g["sha256"][f"final_{n}.bin"], n
    finally:
        e.close()


def test_pool_rejects_overflow(B):
    cases = [B.Case(p) for p in small_cases()]
    e = B.Engine(0)
    try:
        e.alloc_pool(cases[0].ncube - 1, 2)
        e.add_case(cases[0])
        with pytest.raises(B.BCMError, match="no free range"):
            e.add_case(cases[1])
        e.alloc_pool(100, 1)
        e.add_case(cases[0])
        with pytest.raises(B.BCMError, match="slots"):
            e.add_case(cases[2])
    finally:
        e.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_two_level_case_pooled_matches_alone(B, precision):
    # A two-level case between one-size cases in one pool: its two-level
    # faces use pool numbering and touch no other case.
    params = [B.case_params("tiny"), B.case_params("ml_tiny"), B.case_params("tiny", seed=5)]
    alone = []
    for p in params:
        c = B.Case(p)
        e = B.Engine(0)
        try:
            e.alloc_pool(c.ncube - 1, 1, precision)
            s = e.add_case(c)
            er = e.run(2, 2)[:, s, :]
            alone.append(({n: e.case_state(s, n) for n in ("U1_", "U1", "U1q")}, er))
        finally:
            e.close()
    cases = [B.Case(p) for p in params]
    e = B.Engine(0)
    try:
        e.alloc_pool(sum(c.ncube - 1 for c in cases), 3, precision)
        slots = [e.add_case(c) for c in cases]
        er = e.run(2, 2)
        for s, (st, er1) in zip(slots, alone):
            for n, a in st.items():
                bitwise(e.case_state(s, n), a, f"slot {s} {n}")
            bitwise(er[:, s, :], er1, f"slot {s} er")
    finally:
        e.close()
