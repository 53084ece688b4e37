"""GPU tests at the batched workloads' full size: the 64 independent sphere
cases of the batched workload (n^3 cubes with n drawn from {2,3,4,5,6,8}, M in
[0.1, 0.5], Re in [50, 1000], centre offset +-0.1; 7,659 cubes, 31.4M cells)
in one case pool.

Bars:
* FP64 pool (the stage path the benchmark runs: update and face halos pushed
  into the second copy of the iterate): every state array and every norm of
  every case is bitwise the case's run alone, and the smallest (8 cubes) and
  a largest (512 cubes) case are bitwise the FP64 oracle after one physical
  step of 10 inner iterations.  The per-case semantics are the reference's
  time loop (main.cpp:1146-1357) run once per case.
* Mixed pool, the same 64 cases: per case, density relative L2 (fluid cells)
  <= 1e-5 and drag-coefficient relative error <= 1e-4 against the FP64 pool
  after 5 steps; the 8- and 512-cube cases are bitwise their mixed runs alone.
"""
import numpy as np
import pytest

from bcm_cgli_amd.configs import batch_cases
from helpers import describe_mismatch

pytestmark = pytest.mark.gpu

PUSH = (1, 1, 2)         # marching flux, update and face halos pushed to the second copy
N_INNER = 10
RHO_L2_TOL = 1e-5
CD_TOL = 1e-4
FP64_NAMES = ("U1_", "U1", "U1q", "U1p1", "U1p2", "Rku1", "Residual1")


def bitwise(a, b, what):
    if not np.array_equal(a.view(np.int64), b.view(np.int64)):
        raise AssertionError(f"{what}: {describe_mismatch(a, b)}")


@pytest.fixture(scope="module")
def batch(B):
    params = batch_cases(64)
    cases = [B.Case(p, device=0) for p in params]
    assert sum(c.ncube - 1 for c in cases) == 7659
    return params, cases


def pool(B, cases, precision):
    e = B.Engine(0)
    e.alloc_pool(sum(c.ncube - 1 for c in cases), len(cases), precision)
    slots = [e.add_case(c) for c in cases]
    if precision == "fp64":
        e.set_flux(*PUSH)
    return e, slots


def extremes(cases):
    sizes = [c.ncube - 1 for c in cases]
    return sizes.index(min(sizes)), sizes.index(max(sizes))


def test_batch_fp64_pool_bitwise_alone_and_oracle(B, O, batch):
    params, cases = batch
    small, large = extremes(cases)
    assert cases[small].ncube - 1 == 8 and cases[large].ncube - 1 == 512
    e, slots = pool(B, cases, "fp64")
    try:
        # Step 1, then the oracle on the two extreme cases.
        er1 = e.run(1, N_INNER)
        for n in (small, large):
            orc = O.OracleRun(cases[n], nthreads=16)
            her = np.asarray(orc.run(1, N_INNER))
            for name in FP64_NAMES:
                bitwise(e.case_state(slots[n], name), orc.state[name], f"case {n} {name} vs oracle")
            np.testing.assert_allclose(er1[:, slots[n], :5], her[:, :5], rtol=1e-12, atol=0)
            print(f"case {n} ({cases[n].ncube - 1} cubes): bitwise the oracle after 1 step")
        # Step 2; every case against its run alone (same two calls).
        er2 = e.run(1, N_INNER)
        for n, (p, s) in enumerate(zip(params, slots)):
            c = B.Case(p, device=0)
            a = B.Engine(0)
            try:
                a.load_case(c)
                a.set_flux(*PUSH)
                ea1 = a.run(1, N_INNER)
                ea2 = a.run(1, N_INNER)
                for name in FP64_NAMES:
                    bitwise(e.case_state(s, name), a.get_state(name), f"case {n} {name}")
            finally:
                a.close()
            bitwise(er1[:, s, :], ea1, f"case {n} er step 1")
            bitwise(er2[:, s, :], ea2, f"case {n} er step 2")
    finally:
        e.close()


def fluid_rho(case, st):
    fluid = case.array("FWS")[1:, 2:18, 2:18, 2:18] == 1
    return st[1:, 2:18, 2:18, 2:18, 0][fluid]


def test_batch_mixed_pool_within_error_bars(B, batch):
    params, cases = batch
    n_steps = 5
    e, slots = pool(B, cases, "fp64")
    try:
        ref_er = e.run(n_steps, N_INNER)[-1]
        ref_rho = [fluid_rho(c, e.case_state(s, "U1")) for c, s in zip(cases, slots)]
    finally:
        e.close()
    e, slots = pool(B, cases, "mixed")
    try:
        er = e.run(n_steps, N_INNER)[-1]
        worst = (0.0, 0.0)
        for n, (c, s) in enumerate(zip(cases, slots)):
            rho = fluid_rho(c, e.case_state(s, "U1"))
            l2 = float(np.linalg.norm(rho - ref_rho[n]) / np.linalg.norm(ref_rho[n]))
            cd = float(abs(er[s, 5] - ref_er[s, 5]) / abs(ref_er[s, 5]))
            assert np.isfinite(l2) and l2 <= RHO_L2_TOL, (n, l2)
            assert np.isfinite(cd) and cd <= CD_TOL, (n, cd)
            worst = (max(worst[0], l2), max(worst[1], cd))
        print(f"64 cases, {n_steps} steps: worst density rel L2 {worst[0]:.3e}, "
              f"worst Cd rel err {worst[1]:.3e}")
        small, large = extremes(cases)
        for n in (small, large):
            a = B.Engine(0)
            try:
                a.alloc_pool(cases[n].ncube - 1, 1, "mixed")
                sa = a.add_case(B.Case(params[n], device=0))
                ea = a.run(n_steps, N_INNER)[-1]
                for name in ("U1_", "U1", "U1q"):
                    bitwise(e.case_state(slots[n], name), a.case_state(sa, name), f"case {n} {name}")
                bitwise(er[slots[n]], ea[sa], f"case {n} er")
            finally:
                a.close()
    finally:
        e.close()
