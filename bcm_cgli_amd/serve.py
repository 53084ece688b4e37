"""Serving loop: independent flow cases admitted into one device-resident case
pool by HBM budget and run with continuous batching.

Requests arrive over time (each with its own cube count, M, Re and
icosphere level).  A pool of host threads builds each case (grid flags,
IB tables, initial state) as soon as it arrives.  The device loop admits built
cases in arrival order while a case slot and a contiguous range of pool cubes
are free (the pool is sized from the HBM budget once; the engine places a
case in the smallest free range that holds it); a request that does not fit
yet may be passed by later ones that do, at most max_skips times.  It then
runs one physical step of every resident case per iteration, every kernel
launch covering all resident cubes.  After each step a case whose inner
residual fell below the tolerance, or that reached the step limit, is retired
and its cubes are reused by the next admissions; a case whose norms turned
NaN or infinite (diverged) is retired as failed.  So requests arriving while
others run are admitted as soon as slots and memory free up, and their work
overlaps on the device.

Inner residual of a case: max over m = 1..5 of er[m] / scale[m] at the last
inner iteration of a step, with scales (P0, U0, U0, U0, P0/rho0); er[1..5] are
the reference's change norms (BCM_Flux_XYZ_Viscous_Runge_kutta.cpp:2530-2635).
The reference's own inner loop instead stops on running-max-normalised norms
below 0.02 (main.cpp:1248-1271); a serving request needs an absolute stop.
With inner_tol set, each step also runs that inner control: a case's inner
loop ends once its five running-max normalised norms are below inner_tol (or
after n_inner iterations), the maxima carried from step to step per request,
and the residual is taken at the case's last inner iteration of the step.
"""
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Request:
    id: int
    params: dict
    arrival: float               # seconds after the start of run()
    cubes: int = 0
    ready: float = None          # host setup finished
    admit: float = None          # resident on the device
    done: float = None
    steps: int = 0
    converged: bool = False
    residual: float = None
    er: np.ndarray = None        # er[1..7] of the last inner iteration
    E: np.ndarray = None         # running maxima E1..E5 (inner control)
    inner_iters: int = 0         # inner iterations run
    slot: int = None
    case: object = None
    scales: tuple = None
    state: dict = field(default_factory=dict)
    error: str = None
    max_steps: int = None        # per-request step limit (None: the server's)
    skipped: int = 0             # later requests admitted ahead of this one

    @property
    def tts(self):
        return None if self.done is None else self.done - self.arrival


def residual(er7, scales):
    """Inner residual from er[1..7] of one inner iteration."""
    return max(float(er7[m]) / scales[m] for m in range(5))


class Server:
    """Continuous-batching server over one engine.

    engine: an Engine (or an object with the same pool methods).
    make_case: params -> case object with .ncube, .phys, .cells() (B.Case).
    """

    def __init__(self, engine, make_case, precision="mixed", budget_bytes=64 << 30,
                 bytes_per_cube=None, max_cases=64, n_inner=10, max_steps=300, tol=1e-8,
                 prep_threads=4, keep_states=False, clock=time.perf_counter, inner_tol=None,
                 progress=None, max_skips=8, on_step=None):
        self.engine = engine
        self.make_case = make_case
        self.precision = precision
        self.n_inner = int(n_inner)
        self.inner_tol = None if inner_tol is None else float(inner_tol)
        self.progress = progress     # called with (finished, total) as requests finish
        # Called after every step with (steps so far, {slot: request} of the
        # resident requests), on the device loop's thread.
        self.on_step = on_step
        self.max_steps = int(max_steps)
        self.tol = float(tol)
        self.max_cases = int(max_cases)
        self.max_skips = int(max_skips)
        self.keep_states = keep_states
        self.clock = clock
        self.prep_threads = int(prep_threads)
        if bytes_per_cube is None:
            from . import bytes_per_cube as bpc
            bytes_per_cube = bpc(precision)
        self.cap_cubes = int(budget_bytes // bytes_per_cube)
        if self.cap_cubes < 1:
            raise ValueError("HBM budget holds no cube")
        self.budget_bytes = budget_bytes
        self.requests = []
        self._ready = []
        self._cv = threading.Condition()
        self.trace = []   # (time, resident cases, resident cubes) per step

    def submit(self, params, arrival, max_steps=None):
        r = Request(id=len(self.requests), params=dict(params), arrival=float(arrival),
                    max_steps=max_steps)
        self.requests.append(r)
        return r.id

    # -- host side -----------------------------------------------------------
    def _prepare(self, r, t0):
        delay = t0 + r.arrival - self.clock()
        if delay > 0:
            time.sleep(delay)
        try:
            r.case = self.make_case(r.params)
            r.cubes = r.case.ncube - 1
            ph = r.case.phys
            r.scales = (ph["P0"], ph["U0"], ph["U0"], ph["U0"], ph["P0"] / ph["rho0"])
        except Exception as exc:   # reported per request, the loop goes on
            r.error = f"setup: {exc}"
        r.ready = self.clock() - t0
        with self._cv:
            self._ready.append(r)
            self._cv.notify()

    # -- device loop ---------------------------------------------------------
    def _admit(self, pending, resident, t0):
        """Admits pending requests (in arrival order) while a slot and a
        range are free; returns the number of requests finished here (too
        large for the pool, or refused by the engine)."""
        eng = self.engine
        finished = 0
        for r in list(pending):
            if r.cubes > self.cap_cubes:
                pending.remove(r)
                r.error = "case larger than the pool"
                r.done = self.clock() - t0
                finished += 1
                continue
            info = eng.pool_info()
            if info["active"] >= self.max_cases:
                break
            if info["free_range"] < r.cubes:
                # Does not fit now: later requests may pass it, a bounded
                # number of times.
                if r.skipped >= self.max_skips:
                    break
                continue
            try:
                r.slot = eng.add_case(r.case)
            except Exception as exc:   # reported per request, the loop goes on
                pending.remove(r)
                r.error = f"admission: {exc}"
                r.done = self.clock() - t0
                finished += 1
                continue
            r.E = np.zeros(5)
            r.admit = self.clock() - t0
            pending.remove(r)
            resident[r.slot] = r
            for q in pending:
                if q.arrival < r.arrival or (q.arrival == r.arrival and q.id < r.id):
                    q.skipped += 1
        return finished

    def run(self):
        eng = self.engine
        eng.alloc_pool(self.cap_cubes, self.max_cases, self.precision)
        pending, resident = [], {}
        n_total = len(self.requests)
        n_done = 0
        t0 = self.clock()
        cell_iters = 0
        busy = 0.0
        steps = 0
        order = sorted(self.requests, key=lambda r: (r.arrival, r.id))
        pool = ThreadPoolExecutor(max_workers=self.prep_threads)
        futures = [pool.submit(self._prepare, r, t0) for r in order]
        max_resident = 0
        try:
            while n_done < n_total:
                with self._cv:
                    if not self._ready and not resident and not pending:
                        self._cv.wait(timeout=0.05)
                    pending.extend(self._ready)
                    self._ready.clear()
                pending.sort(key=lambda r: (r.arrival, r.id))
                # Setup failures finish at once.
                for r in [r for r in pending if r.error]:
                    pending.remove(r)
                    r.done = self.clock() - t0
                    n_done += 1
                # Admission in arrival order while a slot and a range are free.
                # A request that does not fit the largest free range is passed
                # over by later ones that do, at most max_skips times; after
                # that nothing behind it is admitted until it fits.
                n_done += self._admit(pending, resident, t0)
                if not resident:
                    continue
                max_resident = max(max_resident, len(resident))
                ts = self.clock()
                if self.inner_tol is None:
                    er = eng.run(1, self.n_inner)      # (n_inner, max_cases, 7)
                    n_it = {slot: self.n_inner for slot in resident}
                else:
                    E = np.zeros((eng.max_cases, 5))
                    for slot, r in resident.items():
                        E[slot] = r.E
                    iters, er, E = eng.run_converge(1, self.n_inner, self.inner_tol, E=E)
                    n_it = {slot: int(iters[0, slot]) for slot in resident}
                    for slot, r in resident.items():
                        r.E = np.array(E[slot])
                busy += self.clock() - ts
                steps += 1
                cell_iters += sum(r.case.cells() * n_it[s] for s, r in resident.items())
                self.trace.append((self.clock() - t0, len(resident),
                                   sum(r.cubes for r in resident.values())))
                for slot, r in list(resident.items()):
                    r.steps += 1
                    r.inner_iters += n_it[slot]
                    r.er = np.array(er[n_it[slot] - 1, slot, :])
                    r.residual = residual(r.er, r.scales)
                    # A case whose norms or force sums of any inner iteration
                    # of the step are not finite has diverged (a non-positive
                    # density turns the square roots of the fluxes into NaN
                    # within one stage): it is retired as failed.  Cases share
                    # no data, so the others are unaffected.
                    diverged = not np.all(np.isfinite(er[:n_it[slot], slot, :]))
                    if diverged:
                        r.error = f"diverged: non-finite norms in step {r.steps}"
                    r.converged = not diverged and r.residual < self.tol
                    limit = r.max_steps if r.max_steps is not None else self.max_steps
                    if diverged or r.converged or r.steps >= limit:
                        if self.keep_states:
                            r.state = {n: eng.case_state(slot, n) for n in ("U1_", "U1", "U1q")}
                        eng.remove_case(slot)
                        r.done = self.clock() - t0
                        r.case = r.case if self.keep_states else None
                        del resident[slot]
                        n_done += 1
                        if self.progress:
                            self.progress(n_done, n_total)
                if self.on_step:
                    self.on_step(steps, resident)
        finally:
            pool.shutdown(wait=True)
            for f in futures:
                f.result()
        wall = self.clock() - t0
        ok = [r for r in self.requests if r.error is None]
        tts = np.array([r.tts for r in ok])
        setup = np.array([r.ready - r.arrival for r in ok])
        wait = np.array([r.admit - r.ready for r in ok if r.admit is not None])

        def pct(a, q):
            return float(np.percentile(a, q)) if len(a) else None
        return {
            "requests": n_total,
            "failed": sum(1 for r in self.requests if r.error),
            "diverged": sum(1 for r in self.requests if r.error and r.error.startswith("diverged")),
            "wall_s": wall,
            "device_busy_s": busy,
            "steps": steps,
            "cell_iters": cell_iters,
            "mcell_iter_s": cell_iters / wall / 1e6 if wall > 0 else 0.0,
            "mcell_iter_s_busy": cell_iters / busy / 1e6 if busy > 0 else 0.0,
            "tts_p50_s": pct(tts, 50),
            "tts_p99_s": pct(tts, 99),
            "setup_p50_s": pct(setup, 50),
            "setup_p99_s": pct(setup, 99),
            "admission_wait_p50_s": pct(wait, 50),
            "admission_wait_p99_s": pct(wait, 99),
            "converged": sum(1 for r in self.requests if r.converged),
            "inner_control": ("fixed" if self.inner_tol is None
                              else f"reference, tolerance {self.inner_tol:g}"),
            "inner_iters_per_step": (sum(r.inner_iters for r in ok) /
                                     max(sum(r.steps for r in ok), 1)),
            "max_resident": max_resident,
            "mean_resident": (float(np.mean([n for _, n, _ in self.trace]))
                              if self.trace else 0.0),
            "mean_resident_cubes": (float(np.mean([c for _, _, c in self.trace]))
                                    if self.trace else 0.0),
            "busy_fraction": busy / wall if wall > 0 else 0.0,
            "pool_cubes": self.cap_cubes,
        }


def poisson_requests(n, rate, seed=20260503, cubes=(27, 1000), ico_levels=(3, 5)):
    """Request stream: Poisson arrivals at `rate` per second; per request a
    near-cubic a x b x c lattice with a*b*c uniform in [cubes], M uniform in
    [0.1, 0.5], Re uniform in [50, 1000], icosphere level uniform in
    [ico_levels], sphere of diameter 0.25 box lengths near the box centre."""
    r = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += r.exponential(1.0 / rate)
        target = int(r.integers(cubes[0], cubes[1] + 1))
        a = max(1, int(round(target ** (1 / 3))))
        b = max(1, int(round((target / a) ** 0.5)))
        c = max(1, int(round(target / (a * b))))
        dims = sorted((a, b, c), reverse=True)
        L = 1.0 / dims[0]
        box = [d * L for d in dims]
        centre = tuple(0.5 * bx + r.uniform(-0.1, 0.1) * min(box) for bx in box)
        d = 0.25 * min(box)
        p = dict(lattice=tuple(dims), cube_size=L, origin=(0.0, 0.0, 0.0), sphere_center=centre,
                 sphere_diameter=d, ico_level=int(r.integers(ico_levels[0], ico_levels[1] + 1)),
                 mach=float(r.uniform(0.1, 0.5)), reynolds=float(r.uniform(50, 1000)),
                 seed=int(seed * 1000 + i), noise_amp=1e-3)
        out.append((t, p))
    return out


def measure_capacity(Engine, Case, reqs, precision, n_inner, n_probe=16, steps=2):
    """Batched throughput (cell-iterations per second) of the first requests
    resident together: the capacity the offered load is set against."""
    cases = [Case(p) for _, p in reqs[:n_probe]]
    eng = Engine(0)
    try:
        eng.alloc_pool(sum(c.ncube - 1 for c in cases), len(cases), precision)
        for c in cases:
            eng.add_case(c)
        eng.run(1, n_inner, record_er=False)
        eng.sync()
        t = time.perf_counter()
        eng.run(steps, n_inner, record_er=False)
        eng.sync()
        dt = time.perf_counter() - t
    finally:
        eng.close()
    return sum(c.cells() for c in cases) * n_inner * steps / dt


def serve_run(n_requests, max_steps, n_inner, rate=None, capacity=None, inner_tol=None,
              seed=20260503, log=None):
    """Serves a seeded Poisson stream of the serving workload's requests
    (mixed precision, 64 GiB HBM budget, case setup on 4 host threads with
    the IB search on the GPU).  The arrival rate, unless given, offers 0.8 of
    the batched capacity (cell-iterations per second: given, or measured on
    the stream's first 16 requests) to requests running the full step limit.
    Returns (statistics, configuration)."""
    from . import Case, Engine, bytes_per_cube

    def make_case(p):
        return Case(p, device=0)
    prec = "mixed"
    shape = poisson_requests(n_requests, 1.0, seed=seed)
    mean_cells = float(np.mean([np.prod(p["lattice"]) * 4096 for _, p in shape]))
    if rate is None:
        if capacity is None:
            capacity = measure_capacity(Engine, make_case, shape, prec, n_inner)
        rate = 0.8 * capacity / (mean_cells * n_inner * max_steps)
    reqs = [(t / rate, p) for t, p in shape]
    eng = Engine(0)

    def progress(done, total):
        if log and (done % 8 == 0 or done == total):
            log(f"{done}/{total} requests finished")
    srv = Server(eng, make_case, precision=prec, budget_bytes=64 << 30,
                 bytes_per_cube=bytes_per_cube(prec), max_cases=64, n_inner=n_inner,
                 max_steps=max_steps, tol=1e-8, inner_tol=inner_tol, progress=progress)
    for t, p in reqs:
        srv.submit(p, t)
    if log:
        log(f"serving {len(reqs)} requests at {rate:.3f}/s")
    try:
        st = srv.run()
    finally:
        eng.close()
    inner_desc = (f"{n_inner} inner iterations" if inner_tol is None else
                  f"at most {n_inner} inner iterations (the reference's inner control, tolerance "
                  f"{inner_tol:g})")
    cfg = {
        "workload": f"serving: {len(reqs)} requests, Poisson arrivals at {rate:.3f}/s, cubes "
                    "uniform in [27,1000] (near-cubic lattices), M in [0.1,0.5], Re in "
                    f"[50,1000], icosphere level 3-5, until inner residual < 1e-8 or "
                    f"{max_steps} steps of {inner_desc}, mixed precision, HBM budget 64 GiB, "
                    "case setup on 4 host threads with the IB search on the GPU",
        "requests": len(reqs),
        "rate_per_s": rate,
        "capacity_mcell_iter_s": None if capacity is None else capacity / 1e6,
        "offered_load_at_step_limit": 0.8 if capacity is not None else None,
        "mean_cells_per_request": mean_cells,
    }
    return st, cfg


def bench(args, metric):
    """The serving workload of bench.py: one JSON record."""
    import sys

    def log(m):
        print(f"[bench] {m}", file=sys.stderr, flush=True)
    st, cfg = serve_run(args.requests, args.max_steps, args.inner, rate=args.rate,
                        inner_tol=getattr(args, "inner_tol", None), log=log)
    return {
        "metric": metric,
        "value": round(st["mcell_iter_s"], 3),
        "unit": "Mcell-iter/s",
        "n_gpus": 1,
        "steps": st["steps"],
        "warmup": 0,
        "ms_per_step": round(st["device_busy_s"] / max(st["steps"], 1) * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 stages + f64 time levels",
        "data": "synthetic (generated sphere cases, seeded Poisson request stream)",
        "config": cfg,
        "serving": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in st.items()},
    }
