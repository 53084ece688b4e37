"""Throughput benchmark of the plain FP64 per-iteration solver path on one MI355X.

Workload: the 16x16x8-cube sphere case (8,388,608 cells, level-5 icosphere,
M 0.3, Re 300).  One step = one physical time step = 10 inner iterations of
3 RK stages (domain BCs, IB ghost fill, halo, flux + residual + preconditioned
update, halo) plus the physical-time shift, replayed as one captured HIP graph.
The default 200 steps are the case's full run.

Prints one JSON line.  value = cells x inner iterations / wall seconds / 1e6
(Mcell-iter/s), inputs resident in HBM before the timed region.

--workload batch runs the batched form instead: 64 independent sphere cases
(n^3 cubes, n drawn from {2,3,4,5,6,8}; M, Re and sphere position drawn from a
seed) resident together in one case pool, every kernel launch covering all
their cubes, 10 inner iterations per physical step.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-iter/s + %HBM roofline, mixed vs FP64 at ≤1e-4 Cd err; p99 latency"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# Compulsory HBM traffic of one cell and one RK stage of the flux/update pass:
# read the inner iterate, U1, U1q (40 B each) and the flag (4 B), write the
# stage result (40 B).
ALGO_BYTES_PER_CELL_STAGE = 164


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed physical steps (default: 200 single, 50 batch)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="cfg2")
    ap.add_argument("--workload", default="single", choices=("single", "batch"))
    ap.add_argument("--cases", type=int, default=64)
    ap.add_argument("--inner", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--flux-reps", type=int, default=30)
    return ap.parse_args()


def read_traffic(config):
    """Measured HBM bytes per flux launch from the committed rocprofv3 counter summary."""
    path = os.path.join(ROOT, "profiles", f"flux_{config}_traffic.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def cpu_baseline(case, n_threads):
    """Times the FP64 C oracle (a restatement of the reference CPU path) on a
    bounded sample: one inner iteration (3 stages) of the full case."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    orc = O.OracleRun(case, nthreads=n_threads)
    t0 = time.perf_counter()
    orc.run(1, 1)
    dt = time.perf_counter() - t0
    # orc.run(1, 1) also performs one physical-time shift (negligible).
    return case.cells() / dt / 1e6, dt


def bench_batch(args):
    import bcm_cgli_amd as B
    from bcm_cgli_amd.configs import batch_cases
    params = batch_cases(args.cases)
    t_setup = time.perf_counter()
    cases = [B.Case(p) for p in params]
    setup_s = time.perf_counter() - t_setup
    cubes = sum(c.ncube - 1 for c in cases)
    cells = sum(c.cells() for c in cases)
    print(f"[bench] {len(cases)} cases, {cubes} cubes built in {setup_s:.1f} s", file=sys.stderr,
          flush=True)
    eng = B.Engine(0)
    eng.alloc_pool(cubes, len(cases))
    for c in cases:
        eng.add_case(c)
    n_inner = args.inner
    choice = eng.autotune()
    if args.warmup > 0:
        eng.run(args.warmup, n_inner, record_er=False)
    eng.sync()
    t0 = time.perf_counter()
    eng.run(args.steps, n_inner, record_er=False)
    eng.sync()
    wall = time.perf_counter() - t0
    value = cells * n_inner * args.steps / wall / 1e6
    print(f"[bench] timed {args.steps} steps in {wall:.2f} s", file=sys.stderr, flush=True)
    flux_ms = eng.time_part("flux_update", RK=2, reps=args.flux_reps)
    algo_bytes = ALGO_BYTES_PER_CELL_STAGE * cells
    achieved = algo_bytes / (flux_ms * 1e-3) / 1e9
    sizes = {}
    for p in params:
        sizes[p["lattice"][0] ** 3] = sizes.get(p["lattice"][0] ** 3, 0) + 1
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mcell-iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (generated sphere cases, seeded parameters and density noise)",
        "config": {
            "workload": f"batch: {len(cases)} independent sphere cases in one case pool, "
                        f"cube counts drawn from {{8,27,64,125,216,512}}, M in [0.1,0.5], "
                        f"Re in [50,1000], centre offset +-0.1, {n_inner} inner iterations per step, "
                        f"plain FP64",
            "cases": len(cases),
            "cubes": cubes,
            "cells": cells,
            "cases_per_cube_count": sizes,
            "parallelism": "single GPU, every launch over all cases, hipGraph per physical step",
            "setup_s": round(setup_s, 3),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_flux_march over all cases (one stage)",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "algorithmic_bytes_per_launch": algo_bytes,
            "kernel_ms": round(flux_ms, 4),
        },
        "flux_choice": choice,
    }
    print(json.dumps(out), flush=True)
    eng.close()


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "cases are independent and would shard by cube or by case, see DESIGN.md",
              file=sys.stderr)
        sys.exit(2)
    if args.steps is None:
        args.steps = 50 if args.workload == "batch" else 200
    if args.workload == "batch":
        return bench_batch(args)
    import bcm_cgli_amd as B
    t_setup = time.perf_counter()
    case = B.Case(B.case_params(args.config))
    setup_s = time.perf_counter() - t_setup
    print(f"[bench] case {args.config} built in {setup_s:.1f} s", file=sys.stderr, flush=True)
    eng = B.Engine(0)
    eng.load_case(case)
    cells = case.cells()
    n_inner = args.inner

    eng.autotune_choice = eng.autotune()
    # Warmup (also captures the graph).
    if args.warmup > 0:
        eng.run(args.warmup, n_inner, record_er=False)
    eng.sync()
    t0 = time.perf_counter()
    eng.run(args.steps, n_inner, record_er=False)
    eng.sync()
    wall = time.perf_counter() - t0
    value = cells * n_inner * args.steps / wall / 1e6
    print(f"[bench] timed {args.steps} steps in {wall:.2f} s", file=sys.stderr, flush=True)

    # Dominant kernel: flux + residual + preconditioner (+ the RK update when the
    # tuned choice fuses it), timed with HIP events on the engine stream right
    # after the timed region (same resident state).
    flux_ms = eng.time_part("flux_update", RK=2, reps=args.flux_reps)
    update_ms = eng.time_part("update", RK=2, reps=args.flux_reps)
    halo_ms = eng.time_part("interface", RK=1, reps=args.flux_reps)
    stage_ms = eng.time_part("stage", RK=1, reps=max(3, args.flux_reps // 3))
    algo_bytes = ALGO_BYTES_PER_CELL_STAGE * cells
    achieved = algo_bytes / (flux_ms * 1e-3) / 1e9
    traffic = read_traffic(args.config)
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mcell-iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (generated sphere case, seeded density noise)",
        "config": {
            "workload": f"{args.config}: sphere D=0.25, {case.params['lattice'][0]}x{case.params['lattice'][1]}x"
                        f"{case.params['lattice'][2]} cubes of 16^3 cells, M={case.params['mach']}, "
                        f"Re={case.params['reynolds']}, icosphere level {case.params['ico_level']}, "
                        f"{n_inner} inner iterations per step, plain FP64",
            "cells": cells,
            "cubes": case.ncube - 1,
            "inner_per_step": n_inner,
            "ib_entries": case.info["nbc_minus"] + case.info["nbc_plus"],
            "parallelism": "single GPU, hipGraph per physical step",
            "setup_s": round(setup_s, 3),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_flux_march (flux + residual + preconditioned increment + RK update, one stage)",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": algo_bytes,
            "kernel_ms": round(flux_ms, 4),
        },
        "flux_choice": eng.autotune_choice,
        "kernels_ms": {"flux_update": round(flux_ms, 4), "update_unfused": round(update_ms, 4),
                       "halo_edge": round(halo_ms, 4), "stage": round(stage_ms, 4)},
    }
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline sample", file=sys.stderr, flush=True)
        n_threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        n_threads = max(1, min(n_threads, int(os.environ.get("OMP_NUM_THREADS", n_threads))))
        v, dt = cpu_baseline(case, n_threads)
        out["cpu_baseline"] = {
            "value": round(v, 4),
            "unit": "Mcell-iter/s",
            "cores": n_threads,
            "kind": "port",
            "sample": f"one inner iteration (3 RK stages) of the full {args.config} case, "
                      f"{dt:.1f} s, FP64 C oracle with OpenMP over cubes",
        }
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
