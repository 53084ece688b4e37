"""Host logic of the serving loop, on the CPU: admission by HBM budget in
arrival order, continuous batching, retirement on convergence or step limit,
and the latency statistics.  The device is a stand-in with the pool
interface of the engine (first-fit cube ranges, case slots) whose per-case
residual decays at a rate set by the case; the real engine is exercised by
the GPU tests.
"""
import time

import numpy as np
import pytest

from bcm_cgli_amd.serve import Server, poisson_requests, residual


class FakeCase:
    def __init__(self, params):
        self.ncube = params["cubes"] + 1
        self.decay = params.get("decay", 0.5)
        self.nan_after = params.get("nan_after")   # steps after which the norms turn NaN
        self.steps = 0
        self.phys = {"P0": 1.0, "U0": 1.0, "rho0": 1.0}
        if params.get("fail"):
            raise RuntimeError("bad case")

    def cells(self):
        return (self.ncube - 1) * 4096


class FakeEngine:
    """First-fit pool of cubes with case slots; each step multiplies a case's
    change norms by its decay factor."""

    def __init__(self, step_time=0.0005):
        self.log = []
        self.step_time = step_time

    def alloc_pool(self, cap, max_cases, precision):
        self.cap, self.max_cases = cap, max_cases
        self.owner = [-1] * (cap + 1)
        self.slots = {}

    def pool_info(self):
        best = run = 0
        for c in range(1, self.cap + 1):
            run = run + 1 if self.owner[c] < 0 else 0
            best = max(best, run)
        return {"active": len(self.slots), "free_range": best, "max_cases": self.max_cases}

    def add_case(self, case):
        n = case.ncube - 1
        slot = next(s for s in range(self.max_cases) if s not in self.slots)
        run = 0
        for c in range(1, self.cap + 1):
            run = run + 1 if self.owner[c] < 0 else 0
            if run == n:
                for q in range(c - n + 1, c + 1):
                    self.owner[q] = slot
                self.slots[slot] = [case, 1.0]
                self.log.append(("add", slot, n))
                return slot
        raise RuntimeError("no free range")

    def remove_case(self, slot):
        for c in range(1, self.cap + 1):
            if self.owner[c] == slot:
                self.owner[c] = -1
        del self.slots[slot]
        self.log.append(("remove", slot))

    def run(self, n_steps, n_inner):
        time.sleep(self.step_time)
        er = np.zeros((n_inner, self.max_cases, 7))
        for s, v in self.slots.items():
            v[1] *= v[0].decay
            er[:, s, :5] = v[1]
            v[0].steps += 1
            if v[0].nan_after is not None and v[0].steps > v[0].nan_after:
                er[-1, s, 2] = np.nan
        used = sum(1 for o in self.owner if o >= 0)
        assert used <= self.cap
        return er

    def run_converge(self, n_steps, max_inner, tol, E=None):
        # The reference's inner control over the same decaying norms: per
        # inner iteration the running maxima take the norms, and the loop
        # ends once all five norms are below tol times their maxima.
        time.sleep(self.step_time)
        E = np.array(E, dtype=float).reshape(self.max_cases, 5).copy()
        iters = np.zeros((n_steps, self.max_cases), dtype=np.int32)
        er = np.zeros((n_steps * max_inner, self.max_cases, 7))
        for st in range(n_steps):
            for s, v in self.slots.items():
                for it in range(max_inner):
                    v[1] *= v[0].decay
                    e5 = v[1] * np.arange(1.0, 6.0)
                    er[st * max_inner + it, s, :5] = e5
                    E[s] = np.maximum(E[s], e5)
                    iters[st, s] = it + 1
                    if np.all(e5 < tol * E[s]):
                        break
        return iters, er, E

    def case_state(self, slot, name):
        return np.full(3, slot, dtype=float)


def make_server(eng, cap_cubes, **kw):
    return Server(eng, FakeCase, precision="mixed", budget_bytes=cap_cubes * 1000,
                  bytes_per_cube=1000, **kw)


def test_all_requests_finish_and_fifo_admission():
    # Arrivals far enough apart that each request is prepared (on the host
    # threads) before the next one arrives, however loaded the host is.
    eng = FakeEngine(step_time=0.01)
    srv = make_server(eng, 100, max_cases=8, n_inner=2, max_steps=5, tol=1e-30)
    arrivals = [0.0, 0.02, 0.04, 0.06, 0.08]
    for i, t in enumerate(arrivals):
        srv.submit({"cubes": 30}, t)
    st = srv.run()
    assert st["requests"] == 5 and st["failed"] == 0
    assert all(r.done is not None and r.steps == 5 for r in srv.requests)
    admits = [r.admit for r in srv.requests]
    assert admits == sorted(admits)
    # 100 cubes hold three 30-cube cases at a time.
    assert max(n for _, n, _ in srv.trace) == 3
    assert all(c <= 100 for _, _, c in srv.trace)
    assert st["tts_p99_s"] >= st["tts_p50_s"] > 0


def test_continuous_batching_overlaps_requests():
    """A request arriving while another runs is admitted before that one ends."""
    eng = FakeEngine()
    srv = make_server(eng, 100, max_cases=4, n_inner=1, max_steps=200, tol=1e-30)
    srv.submit({"cubes": 10}, 0.0)
    srv.submit({"cubes": 10}, 0.01)
    srv.run()
    a, b = srv.requests
    assert b.admit < a.done
    assert b.admit >= b.arrival


def test_retire_on_convergence_and_reuse_of_cubes():
    eng = FakeEngine()
    srv = make_server(eng, 50, max_cases=4, n_inner=1, max_steps=100, tol=1e-3)
    srv.submit({"cubes": 40, "decay": 0.05}, 0.0)    # converges after 3 steps
    srv.submit({"cubes": 40, "decay": 0.5}, 0.0)     # waits for the first one's cubes
    st = srv.run()
    a, b = srv.requests
    assert a.converged and a.steps == 3
    assert b.converged and b.steps == 10
    assert b.admit >= a.done
    assert st["converged"] == 2
    assert [x[0] for x in eng.log] == ["add", "remove", "add", "remove"]


def test_setup_failure_and_oversized_request_do_not_stop_the_loop():
    eng = FakeEngine()
    srv = make_server(eng, 20, max_cases=4, n_inner=1, max_steps=2, tol=1e-30)
    srv.submit({"cubes": 5, "fail": True}, 0.0)
    srv.submit({"cubes": 50}, 0.0)
    srv.submit({"cubes": 5}, 0.0)
    st = srv.run()
    r0, r1, r2 = srv.requests
    assert r0.error and r1.error and r2.error is None and r2.steps == 2
    assert st["failed"] == 2


def test_slot_limit():
    eng = FakeEngine()
    srv = make_server(eng, 1000, max_cases=2, n_inner=1, max_steps=3, tol=1e-30)
    for _ in range(5):
        srv.submit({"cubes": 1}, 0.0)
    srv.run()
    assert max(n for _, n, _ in srv.trace) == 2


def test_diverged_case_is_quarantined():
    # The second request's norms turn NaN in its third step: it is retired
    # at once as failed, its slot and cubes go to the waiting request, and
    # the others run to their limits.
    eng = FakeEngine()
    srv = make_server(eng, 60, max_cases=2, n_inner=2, max_steps=6, tol=1e-30)
    srv.submit({"cubes": 20}, 0.0)
    srv.submit({"cubes": 20, "nan_after": 2}, 0.0)
    srv.submit({"cubes": 20}, 0.0)
    st = srv.run()
    a, b, c = srv.requests
    assert b.error.startswith("diverged") and b.steps == 3 and not b.converged
    assert a.error is None and a.steps == 6 and c.error is None and c.steps == 6
    assert c.admit >= b.done and c.admit < a.done
    assert st["failed"] == 1 and st["diverged"] == 1
    assert ("remove", b.slot) in eng.log


def test_admission_passes_a_request_that_does_not_fit_a_bounded_number_of_times():
    def run(max_skips):
        # Steps long enough that every request is prepared (on the host
        # threads) well before A's four steps end.
        eng = FakeEngine(step_time=0.01)
        srv = make_server(eng, 100, max_cases=8, n_inner=1, max_steps=4, tol=1e-30,
                          max_skips=max_skips)
        srv.submit({"cubes": 50}, 0.0)      # A
        srv.submit({"cubes": 80}, 0.0)      # B: waits for A's (and C's) cubes
        srv.submit({"cubes": 20}, 0.0)      # C
        srv.submit({"cubes": 20}, 0.0)      # D
        srv.run()
        return srv.requests
    a, b, c, d = run(1)
    assert c.admit < b.admit and c.admit < a.done     # C passes B once
    assert d.admit >= b.admit                          # D may not: B's bound is reached
    assert b.skipped == 1 and b.steps == 4
    a, b, c, d = run(8)
    assert c.admit < b.admit and d.admit < b.admit     # both pass B
    assert b.skipped == 2


def alone_inner_control(decay, n_inner, inner_tol, tol, max_steps):
    """One request by itself under the inner control: (steps, inner
    iterations, last norms, E), written out from the rule directly."""
    v, E, steps, total, last = 1.0, np.zeros(5), 0, 0, None
    while True:
        for it in range(n_inner):
            v *= decay
            last = v * np.arange(1.0, 6.0)
            E = np.maximum(E, last)
            total += 1
            if np.all(last < inner_tol * E):
                break
        steps += 1
        if max(last) < tol or steps >= max_steps:
            return steps, total, last, E


def test_inner_control_per_request():
    # Four requests with different decay rates share the device; each ends
    # its inner loops, carries its maxima and retires as it would alone, and
    # the throughput counts the inner iterations that ran.
    eng = FakeEngine()
    srv = make_server(eng, 60, max_cases=3, n_inner=8, max_steps=12, tol=1e-6,
                      inner_tol=0.05)
    decays = [0.5, 0.7, 0.9, 0.3]
    for i, d in enumerate(decays):
        srv.submit({"cubes": 20, "decay": d}, 0.001 * i)
    st = srv.run()
    assert st["failed"] == 0
    assert st["inner_control"] == "reference, tolerance 0.05"
    total = 0
    for r, d in zip(srv.requests, decays):
        steps, iters, last, E = alone_inner_control(d, 8, 0.05, 1e-6, 12)
        assert (r.steps, r.inner_iters) == (steps, iters), d
        assert np.allclose(r.er[:5], last, rtol=1e-12) and np.allclose(r.E, E, rtol=1e-12)
        total += iters * 20 * 4096
    # Some inner loops ended before the cap.
    assert any(r.inner_iters < 8 * r.steps for r in srv.requests)
    assert st["cell_iters"] == total
    assert st["inner_iters_per_step"] == pytest.approx(
        sum(r.inner_iters for r in srv.requests) / sum(r.steps for r in srv.requests))


def test_residual_scaling():
    er = np.array([2.0, 3.0, 1.0, 0.5, 8.0, 0.0, 0.0])
    assert residual(er, (2.0, 1.0, 1.0, 1.0, 4.0)) == 3.0


def test_poisson_stream_shape():
    reqs = poisson_requests(400, rate=5.0, seed=1)
    t = np.array([a for a, _ in reqs])
    assert np.all(np.diff(t) > 0)
    assert abs(len(t) / t[-1] - 5.0) < 1.0
    cubes = np.array([np.prod(p["lattice"]) for _, p in reqs])
    assert cubes.min() >= 20 and cubes.max() <= 1100
    for _, p in reqs:
        assert 0.1 <= p["mach"] <= 0.5 and 50 <= p["reynolds"] <= 1000
        assert 3 <= p["ico_level"] <= 5
        a, b, c = p["lattice"]
        box = np.array([a, b, c]) * p["cube_size"]
        ctr = np.array(p["sphere_center"])
        assert np.all(ctr - p["sphere_diameter"] / 2 > 0) and np.all(ctr + p["sphere_diameter"] / 2 < box)


@pytest.mark.slow
def test_poisson_cases_build():
    """A few drawn requests build on the host (flags, IB tables)."""
    import bcm_cgli_amd as B
    for _, p in poisson_requests(3, rate=1.0, seed=3, cubes=(27, 80)):
        c = B.Case(p)
        assert c.info["nbc_minus"] + c.info["nbc_plus"] > 0
