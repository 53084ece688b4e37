"""Serving loop on the GPU: requests arriving while others run are admitted
into the resident case pool and overlap on the device; each request's final
state and norms are bitwise the ones the same case gives when run alone.
"""
import numpy as np
import pytest

from bcm_cgli_amd.serve import Server, poisson_requests

pytestmark = pytest.mark.gpu


def run_alone(B, params, n_steps, n_inner, precision):
    case = B.Case(params)
    e = B.Engine(0)
    try:
        e.alloc_pool(case.ncube - 1, 1, precision)
        s = e.add_case(case)
        er = e.run(n_steps, n_inner)[:, s, :]
        return {n: e.case_state(s, n) for n in ("U1_", "U1", "U1q")}, er[-1]
    finally:
        e.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_served_results_match_alone(B, precision):
    reqs = poisson_requests(5, rate=40.0, seed=11, cubes=(27, 64), ico_levels=(3, 3))
    steps = [6, 2, 3, 2, 4]
    e = B.Engine(0)
    srv = Server(e, B.Case, precision=precision, budget_bytes=150 * B.bytes_per_cube(precision),
                 max_cases=2, n_inner=2, max_steps=100, tol=1e-30, keep_states=True)
    for (t, p), n in zip(reqs, steps):
        srv.submit(p, 0.0, max_steps=n)
    st = srv.run()
    e.close()
    assert st["failed"] == 0 and st["requests"] == 5
    # Budget and slots bind: never more than 2 cases or 150 cubes resident.
    assert max(n for _, n, _ in srv.trace) <= 2
    assert max(c for _, _, c in srv.trace) <= 150
    # Requests overlap on the device: one is admitted while another runs.
    rs = srv.requests
    assert any(a.admit < b.admit < a.done for a in rs for b in rs if a is not b)
    for r, n in zip(rs, steps):
        assert r.steps == n
        st1, er1 = run_alone(B, r.params, n, 2, precision)
        for name in ("U1_", "U1", "U1q"):
            assert np.array_equal(r.state[name], st1[name]), (r.id, name)
        assert np.array_equal(r.er, er1[:7]), r.id


def test_served_serving_distribution_matches_alone(B):
    # Ten requests drawn from the serving workload's own distribution (cube
    # counts uniform in [27, 1000], icosphere level 3-5, M in [0.1, 0.5], Re
    # in [50, 1000]), mixed precision, 10 inner iterations per step, the 1e-8
    # residual stop, per-request step limits of 2-7 steps.  A 2,500-cube
    # budget and 4 slots make later requests wait for retirements and land in
    # ranges earlier requests freed.  Each result is bitwise its run alone.
    reqs = poisson_requests(10, rate=50.0, seed=2026)
    steps = [2 + (i * 3) % 6 for i in range(len(reqs))]
    cubes = [int(np.prod(p["lattice"])) for _, p in reqs]
    assert max(cubes) > 600 and min(cubes) < 400
    assert len({p["ico_level"] for _, p in reqs}) >= 2
    e = B.Engine(0)
    srv = Server(e, lambda p: B.Case(p, device=0), precision="mixed",
                 budget_bytes=2500 * B.bytes_per_cube("mixed"), max_cases=4, n_inner=10,
                 max_steps=300, tol=1e-8, keep_states=True)
    for (t, p), n in zip(reqs, steps):
        srv.submit(p, 0.0, max_steps=n)
    st = srv.run()
    e.close()
    assert st["failed"] == 0 and st["requests"] == len(reqs)
    assert max(c for _, _, c in srv.trace) <= 2500
    rs = srv.requests
    assert any(a.admit < b.admit < a.done for a in rs for b in rs if a is not b)
    for r in rs:
        assert 1 <= r.steps
        st1, er1 = run_alone(B, r.params, r.steps, 10, "mixed")
        for name in ("U1_", "U1", "U1q"):
            assert np.array_equal(r.state[name], st1[name]), (r.id, name)
        assert np.array_equal(r.er, er1[:7]), r.id


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_diverged_request_is_quarantined(B, precision):
    # After its second step, one resident request's iterate gets a NaN
    # density in one fluid cell (uploaded from the host).  Its norms turn NaN
    # in the next step: it is reported failed and its slot freed, while the
    # requests sharing the launches finish bitwise as they do alone.
    reqs = poisson_requests(4, rate=40.0, seed=21, cubes=(27, 64), ico_levels=(3, 3))
    steps = [5, 6, 4, 5]
    target = 1
    e = B.Engine(0)

    def poison(n_steps, resident):
        for slot, r in resident.items():
            if r.id == target and r.steps == 2:
                u = e.case_state(slot, "U1_")
                fluid = np.argwhere(r.case.array("FWS")[1:, 2:18, 2:18, 2:18] == 1)[0]
                c, i, j, k = fluid + np.array([1, 2, 2, 2])
                u[c, i, j, k, 0] = np.nan
                e.set_case_state(slot, "U1_", u)

    srv = Server(e, B.Case, precision=precision, budget_bytes=200 * B.bytes_per_cube(precision),
                 max_cases=3, n_inner=3, max_steps=100, tol=1e-30, keep_states=True,
                 on_step=poison)
    for (t, p), n in zip(reqs, steps):
        srv.submit(p, 0.0, max_steps=n)
    st = srv.run()
    e.close()
    rs = srv.requests
    bad = rs[target]
    assert bad.error is not None and bad.error.startswith("diverged"), bad.error
    assert bad.steps == 3 and not bad.converged
    assert st["failed"] == 1 and st["diverged"] == 1
    for r, n in zip(rs, steps):
        if r.id == target:
            continue
        assert r.error is None and r.steps == n
        st1, er1 = run_alone(B, r.params, n, 3, precision)
        for name in ("U1_", "U1", "U1q"):
            assert np.array_equal(r.state[name], st1[name]), (r.id, name)
        assert np.array_equal(r.er, er1[:7]), r.id


def run_alone_converge(B, params, n_steps, n_inner, inner_tol, precision):
    case = B.Case(params)
    e = B.Engine(0)
    try:
        e.alloc_pool(case.ncube - 1, 1, precision)
        s = e.add_case(case)
        iters, er, E = e.run_converge(n_steps, n_inner, inner_tol)
        last = er[(n_steps - 1) * n_inner + iters[-1, s] - 1, s]
        return ({n: e.case_state(s, n) for n in ("U1_", "U1", "U1q")}, int(iters[:, s].sum()),
                last, E[s])
    finally:
        e.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_served_inner_control_matches_alone(B, precision):
    # Under the reference's inner control the 64-cube requests end inner
    # loops early from step 4 on (tolerance 0.97, at most 8 iterations; the
    # FP64 oracle takes 8, 8, 8, 6, 5, 4 for the first), while the request
    # admitted later still runs full loops in the same launches.  Each result
    # is bitwise the request's run alone, with its maxima carried per step.
    reqs = poisson_requests(3, rate=40.0, seed=11, cubes=(27, 64), ico_levels=(3, 3))
    steps = [6, 3, 6]
    e = B.Engine(0)
    srv = Server(e, B.Case, precision=precision, budget_bytes=150 * B.bytes_per_cube(precision),
                 max_cases=2, n_inner=8, max_steps=100, tol=1e-30, keep_states=True,
                 inner_tol=0.97)
    for (t, p), n in zip(reqs, steps):
        srv.submit(p, 0.0, max_steps=n)
    st = srv.run()
    e.close()
    assert st["failed"] == 0 and st["requests"] == 3
    rs = srv.requests
    assert any(a.admit < b.admit < a.done for a in rs for b in rs if a is not b)
    assert any(r.inner_iters < 8 * r.steps for r in rs)
    assert st["cell_iters"] == sum(r.case.cells() * r.inner_iters for r in rs)
    for r, n in zip(rs, steps):
        assert r.steps == n
        st1, it1, er1, E1 = run_alone_converge(B, r.params, n, 8, 0.97, precision)
        assert r.inner_iters == it1, r.id
        for name in ("U1_", "U1", "U1q"):
            assert np.array_equal(r.state[name], st1[name]), (r.id, name)
        assert np.array_equal(r.er, er1[:7]), r.id
        assert np.array_equal(r.E, E1), r.id


def test_serve_stats(B):
    reqs = poisson_requests(4, rate=20.0, seed=5, cubes=(27, 40), ico_levels=(3, 4))
    e = B.Engine(0)
    srv = Server(e, B.Case, precision="mixed", budget_bytes=1 << 30, max_cases=8, n_inner=2,
                 max_steps=3, tol=1e-8)
    for t, p in reqs:
        srv.submit(p, t)
    st = srv.run()
    e.close()
    assert st["requests"] == 4 and st["failed"] == 0
    assert st["mcell_iter_s"] > 0 and st["tts_p99_s"] >= st["tts_p50_s"] > 0
    for r in srv.requests:
        assert r.arrival <= r.ready <= r.admit <= r.done
        assert np.all(np.isfinite(r.er))
