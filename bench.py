"""Throughput benchmark of the per-iteration solver path on one MI355X.

Headline workload (value): the 16x16x8-cube sphere case (8,388,608 cells,
level-5 icosphere, M 0.3, Re 300) in plain FP64.  One step = one physical time
step = 10 inner iterations of 3 RK stages (domain BCs, IB ghost fill, halo,
flux + residual + preconditioned update, halo) plus the physical-time shift,
replayed as one captured HIP graph.  The default 200 steps are the case's
full run.

value = cells x inner iterations / wall seconds / 1e6 (Mcell-iter/s), inputs
resident in HBM before the timed region.

The same line also reports, under "forms":
  mixed_cfg2   the same case and steps in mixed precision, with the relative
               L2 error of density (fluid cells) and the relative error of the
               drag coefficient (normalised x-force of the last inner
               iteration) against the FP64 run, which is bitwise the FP64
               oracle;
  batch_fp64 / batch_mixed
               64 independent sphere cases (n^3 cubes, n drawn from
               {2,3,4,5,6,8}; M, Re and sphere position drawn from a seed) in
               one case pool, 50 steps of 10 inner iterations, every launch
               covering all cases; the mixed form with per-case errors
               against the FP64 form;
  serve_cfg5   the serving workload on a bounded sample: 48 Poisson requests
               (cube counts uniform in [27, 1000], icosphere level 3-5, M and
               Re as above, mixed precision) until the inner residual is below
               1e-8 or 300 steps, arriving at a rate that offers 0.8 of the
               batch_mixed throughput to requests running all 300 steps;
               p50/p99 time-to-solution, sustained and busy Mcell-iter/s.
--workload serve runs the serving loop alone (see bcm_cgli_amd/serve.py).
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-iter/s + %HBM roofline, mixed vs FP64 at ≤1e-4 Cd err; p99 latency"
HBM_PEAK_GBS = 8000.0
# Compulsory HBM traffic of one cell and one RK stage of the flux/update pass:
# FP64: read the inner iterate (40 B), U1 and U1q (40 B each) and the flag
# (1 B), write the stage result (40 B).  Mixed: read the FP32 iterate (20 B),
# the FP64 dual-time source constant 3 Uref - 4 U1 + U1q formed once per step
# (40 B) and the flag (1 B), write the FP32 stage result (20 B).
ALGO_BYTES = {"fp64": 161, "mixed": 81}
FLUX_KERNEL = {"fp64": "k_flux_march", "mixed": "k_flux_mixed"}   # traffic file stems


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed physical steps")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--inner", type=int, default=10)
    ap.add_argument("--config", default="cfg2")
    ap.add_argument("--precision", default="fp64", choices=("fp64", "mixed"))
    ap.add_argument("--workload", default="single", choices=("single", "serve"))
    ap.add_argument("--no-forms", action="store_true", help="skip the mixed and batched forms")
    ap.add_argument("--batch-cases", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--flux-reps", type=int, default=30)
    ap.add_argument("--mixed-nseg", type=int, default=None, choices=(1, 2, 4),
                    help="pin the mixed flux kernel's runs per cube (default: autotuned)")
    ap.add_argument("--serve-requests", type=int, default=48,
                    help="requests of the serving form (0: skip it)")
    ap.add_argument("--serve-max-steps", type=int, default=300)
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--rate", type=float, default=None, help="serve: arrivals per second")
    ap.add_argument("--max-steps", type=int, default=300)
    ap.add_argument("--inner-tol", type=float, default=None,
                    help="serve: end each request's inner loop as the reference does, once its "
                         "running-max normalised norms fall below this tolerance (0.02 in the "
                         "reference); default: always --inner iterations")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_traffic(kernel, config):
    """Measured HBM bytes per launch from a committed rocprofv3 counter summary."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{config}_traffic.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


# Vector issue rate of one MI355X: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
# (FP64 and unpacked FP32 instructions issue at this rate).
VALU_PEAK_TLANE = 256 * 4 * 16 * 2.4e9 / 1e12


def read_valu(kernel):
    """VALU instructions per cell-stage of a flux kernel from the committed
    SQ counter summary (SQ_INSTS_VALU per wave x 64 lanes / 1,024 cell-stages)."""
    path = os.path.join(ROOT, "profiles", "flux_cfg2_counters.txt")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    at = text.find(kernel)
    if at < 0:
        return None
    m = re.search(r"VALU/wave\s+\d+\s+\((\d+) per cell-stage\)", text[at:])
    return int(m.group(1)) if m else None


# Runs per cube of the mixed flux kernel when pinned (--mixed-nseg); None: autotuned.
MIXED_NSEG = None


def run_pool(B, cases, precision, steps, warmup, n_inner, flux_reps, keep_state):
    """Loads cases into one pool, autotunes, warms up, times `steps` steps.
    Returns timing, the last inner iteration's er per case and (keep_state)
    each case's fluid-cell density after warmup + steps steps."""
    eng = B.Engine(0)
    eng.alloc_pool(sum(c.ncube - 1 for c in cases), len(cases), precision)
    slots = [eng.add_case(c) for c in cases]
    choice = eng.autotune()
    if precision == "mixed" and MIXED_NSEG is not None:
        eng.set_flux(2, MIXED_NSEG, 0)
        choice = {"variant": 2, "nseg": MIXED_NSEG, "fuse": 2, "pinned": True}
    if warmup > 0:
        eng.run(warmup, n_inner, record_er=False)
    eng.sync()
    t0 = time.perf_counter()
    eng.run(steps, n_inner, record_er=False)
    eng.sync()
    wall = time.perf_counter() - t0
    cells = sum(c.cells() for c in cases)
    res = {
        "value": cells * n_inner * steps / wall / 1e6,
        "wall": wall,
        "cells": cells,
        "choice": choice,
        "er": [eng.case_er(s, n_inner - 1)[1:] for s in slots],
    }
    if keep_state:
        res["rho"] = []
        for s, c in zip(slots, cases):
            fluid = c.array("FWS")[1:, 2:18, 2:18, 2:18] == 1
            res["rho"].append(eng.case_state(s, "U1")[1:, 2:18, 2:18, 2:18, 0][fluid])
    # The flux kernel the stages run: with the halo push (fuse 2) or not.
    flux_part = "flux_push" if choice.get("fuse") == 2 else "flux_update"
    res["flux_ms"] = eng.time_part(flux_part, RK=2, reps=flux_reps)
    res["flux_part"] = flux_part
    res["stage_ms"] = eng.time_part("stage", RK=1, reps=max(3, flux_reps // 3))
    res["parts_ms"] = {p: round(eng.time_part(p, RK=1, reps=flux_reps), 4)
                       for p in ("bc", "ghost", "halo_refresh", "interface")}
    if precision == "fp64":
        res["parts_ms"]["update_unfused"] = round(eng.time_part("update", RK=2, reps=flux_reps), 4)
    eng.close()
    return res


def errors(mixed, ref):
    """Per case: relative L2 error of fluid-cell density, relative Cd error."""
    out = []
    for rm, rr, em, er in zip(mixed["rho"], ref["rho"], mixed["er"], ref["er"]):
        l2 = float(np.linalg.norm(rm - rr) / np.linalg.norm(rr))
        cd = float(abs(em[5] - er[5]) / abs(er[5]))
        out.append((l2, cd))
    return out


def roofline(precision, cells, flux_ms, config, push=True):
    algo = ALGO_BYTES[precision] * cells
    achieved = algo / (flux_ms * 1e-3) / 1e9
    form = "writing the update and face halos to the second copy" if push else "in place"
    return {
        "bound": "hbm",
        "kernel": f"{FLUX_KERNEL[precision]} (flux + residual + preconditioned increment + RK "
                  f"update of one stage, all cubes, {form})",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": read_traffic(FLUX_KERNEL[precision], config),
        "algorithmic_bytes_per_launch": algo,
        "kernel_ms": round(flux_ms, 4),
        "valu": valu_roofline(precision, cells, flux_ms),
    }


def valu_roofline(precision, cells, flux_ms):
    """The flux kernels are bound by instruction issue, not HBM: the vector
    instruction rate they reach against the MI355X issue rate."""
    kern = {"fp64": "k_flux_march<true>", "mixed": "k_flux_mixed"}[precision]
    per_cell = read_valu(kern)
    if per_cell is None:
        return None
    achieved = per_cell * cells / (flux_ms * 1e-3) / 1e12
    # HBM fraction the kernel would reach at the full issue rate: its
    # instruction count caps the cell-stage rate, and with it the bytes moved.
    hbm_cap = VALU_PEAK_TLANE * 1e12 / per_cell * ALGO_BYTES[precision] / 1e9 / HBM_PEAK_GBS
    return {
        "instr_per_cell_stage": per_cell,
        "achieved": round(achieved, 3),
        "peak": round(VALU_PEAK_TLANE, 3),
        "unit": "T lane-instr/s",
        "frac": round(achieved / VALU_PEAK_TLANE, 4),
        "hbm_frac_at_full_issue": round(hbm_cap, 4),
        "source": "profiles/flux_cfg2_counters.txt (SQ_INSTS_VALU)",
    }


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


def cpu_baseline_reference(case, params, n_threads, config):
    """Times the reference solver itself, compiled from its own sources by
    oracle/build_ref.sh for this case's cube count (present when the build
    ran where the reference tree is): one inner iteration (3 stages) of the
    full case, the time loop alone as the binary reports it."""
    exe = os.path.join(ROOT, "oracle", "_ref", config, "build", "ref_driver")
    if not os.access(exe, os.X_OK):
        return None
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        case.write_inputs(d)
        env = dict(os.environ, OMP_NUM_THREADS=str(n_threads))
        t0 = time.perf_counter()
        try:
            r = subprocess.run([exe, "1", "1", str(params["seed"]), repr(params["noise_amp"]), "-"],
                               cwd=d, env=env, capture_output=True, text=True, timeout=600)
        except (OSError, subprocess.TimeoutExpired):
            return None
        total = time.perf_counter() - t0
    m = re.search(r"loop_seconds\s+([0-9.]+)", r.stdout)
    if r.returncode != 0 or not m:
        return None
    loop = float(m.group(1))
    return case.cells() / loop / 1e6, loop, total


def main():
    args = parse_args()
    global MIXED_NSEG
    MIXED_NSEG = args.mixed_nseg
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "cases are independent and would shard by cube or by case",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "serve":
        from bcm_cgli_amd import serve
        print(json.dumps(serve.bench(args, METRIC)), flush=True)
        return
    import bcm_cgli_amd as B
    from bcm_cgli_amd.configs import batch_cases
    t_setup = time.perf_counter()
    case = B.Case(B.case_params(args.config), device=0)
    setup_s = time.perf_counter() - t_setup
    log(f"case {args.config} built in {setup_s:.1f} s")
    n_inner = args.inner
    forms = not args.no_forms
    main_prec = args.precision
    r = run_pool(B, [case], main_prec, args.steps, args.warmup, n_inner, args.flux_reps,
                 keep_state=forms)
    log(f"{main_prec}: timed {args.steps} steps in {r['wall']:.2f} s, {r['value']:.1f} Mcell-iter/s")
    cells = case.cells()
    p = case.params
    out = {
        "metric": METRIC,
        "value": round(r["value"], 3),
        "unit": "Mcell-iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["wall"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_prec == "fp64" else "f32 stages + f64 time levels",
        "data": "synthetic (generated sphere case, seeded density noise)",
        "config": {
            "workload": f"{args.config}: sphere D=0.25, {p['lattice'][0]}x{p['lattice'][1]}x"
                        f"{p['lattice'][2]} cubes of 16^3 cells, M={p['mach']}, Re={p['reynolds']}, "
                        f"icosphere level {p['ico_level']}, {n_inner} inner iterations per step, "
                        f"{'plain FP64' if main_prec == 'fp64' else 'mixed precision'}",
            "cells": cells,
            "cubes": case.ncube - 1,
            "inner_per_step": n_inner,
            "ib_entries": case.info["nbc_minus"] + case.info["nbc_plus"],
            "parallelism": "single GPU, hipGraph per physical step",
            "setup_s": round(setup_s, 3),
        },
        "roofline": roofline(main_prec, cells, r["flux_ms"], args.config,
                             r["flux_part"] == "flux_push" or main_prec == "mixed"),
        "flux_choice": r["choice"],
        "kernels_ms": {"flux_update": round(r["flux_ms"], 4), "stage": round(r["stage_ms"], 4),
                       **r["parts_ms"]},
    }
    if forms:
        f = {}
        other = "mixed" if main_prec == "fp64" else "fp64"
        rm = run_pool(B, [case], other, args.steps, args.warmup, n_inner, args.flux_reps, True)
        log(f"{other}: {rm['value']:.1f} Mcell-iter/s")
        ref, mix = (r, rm) if main_prec == "fp64" else (rm, r)
        (l2, cd), = errors(mix, ref)
        f["mixed_cfg2"] = {
            "value": round(mix["value"], 3), "unit": "Mcell-iter/s",
            "speedup_vs_fp64": round(mix["value"] / ref["value"], 3),
            "density_rel_l2": l2, "cd_rel_err": cd,
            "cd_fp64": B.drag_coefficient(ref["er"][0][5], p["sphere_diameter"]),
            "cd_mixed": B.drag_coefficient(mix["er"][0][5], p["sphere_diameter"]),
            "er6_fp64": float(ref["er"][0][5]), "er6_mixed": float(mix["er"][0][5]),
            "cd_definition": "Cd = -er[6] * 2.602 / (pi D^2 / 4), er[6] the normalised x-force "
                             "of the last inner iteration",
            "steps_compared": args.warmup + args.steps,
            "roofline": roofline("mixed", cells, mix["flux_ms"], args.config),
            "flux_choice": mix["choice"],
            "kernels_ms": {"flux_update": round(mix["flux_ms"], 4),
                           "stage": round(mix["stage_ms"], 4), **mix["parts_ms"]},
        }
        params = batch_cases(args.batch_cases)
        t0 = time.perf_counter()
        bcases = [B.Case(q, device=0) for q in params]
        bsetup = time.perf_counter() - t0
        log(f"batch: {len(bcases)} cases built in {bsetup:.1f} s")
        b64 = run_pool(B, bcases, "fp64", args.batch_steps, 1, n_inner, 10, True)
        log(f"batch fp64: {b64['value']:.1f} Mcell-iter/s")
        bmx = run_pool(B, bcases, "mixed", args.batch_steps, 1, n_inner, 10, True)
        log(f"batch mixed: {bmx['value']:.1f} Mcell-iter/s")
        errs = errors(bmx, b64)
        sizes = {}
        for q in params:
            sizes[q["lattice"][0] ** 3] = sizes.get(q["lattice"][0] ** 3, 0) + 1
        batch_cfg = {
            "workload": f"{len(bcases)} independent sphere cases in one case pool, cube counts "
                        "drawn from {8,27,64,125,216,512}, M in [0.1,0.5], Re in [50,1000], "
                        f"centre offset +-0.1, {args.batch_steps} steps of {n_inner} inner iterations",
            "cases": len(bcases), "cubes": sum(c.ncube - 1 for c in bcases),
            "cells": b64["cells"], "cases_per_cube_count": sizes, "setup_s": round(bsetup, 3),
        }
        f["batch_fp64"] = {
            "value": round(b64["value"], 3), "unit": "Mcell-iter/s", "config": batch_cfg,
            "roofline": roofline("fp64", b64["cells"], b64["flux_ms"], "batch",
                                 b64["flux_part"] == "flux_push"),
            "flux_choice": b64["choice"],
        }
        f["batch_mixed"] = {
            "value": round(bmx["value"], 3), "unit": "Mcell-iter/s",
            "speedup_vs_plain_fp64_cfg2": round(bmx["value"] / ref["value"], 3),
            "speedup_vs_batch_fp64": round(bmx["value"] / b64["value"], 3),
            "density_rel_l2_max": max(e[0] for e in errs),
            "cd_rel_err_max": max(e[1] for e in errs),
            "density_rel_l2_per_case": [float(f"{e[0]:.3e}") for e in errs],
            "cd_rel_err_per_case": [float(f"{e[1]:.3e}") for e in errs],
            "cd_fp64_per_case": [round(B.drag_coefficient(e[5], q["sphere_diameter"]), 6)
                                 for e, q in zip(b64["er"], params)],
            "roofline": roofline("mixed", bmx["cells"], bmx["flux_ms"], "batch"),
            "flux_choice": bmx["choice"],
        }
        if args.serve_requests > 0:
            from bcm_cgli_amd import serve
            st, scfg = serve.serve_run(args.serve_requests, args.serve_max_steps, n_inner,
                                       capacity=bmx["value"] * 1e6, log=log)
            rnd = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in st.items()}
            f["serve_cfg5"] = {
                "value": round(st["mcell_iter_s"], 3), "unit": "Mcell-iter/s",
                "value_busy": round(st["mcell_iter_s_busy"], 3),
                "busy_speedup_vs_plain_fp64_cfg2": round(st["mcell_iter_s_busy"] / ref["value"], 3),
                "tts_p50_s": rnd["tts_p50_s"], "tts_p99_s": rnd["tts_p99_s"],
                "config": {**scfg, "capacity_source": "batch_mixed value (64-case pool)"},
                "serving": rnd,
            }
        out["forms"] = f
    if not args.no_cpu_baseline:
        log("CPU baseline sample")
        n_threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        n_threads = max(1, min(n_threads, int(os.environ.get("OMP_NUM_THREADS", n_threads))))
        v, dt = cpu_baseline(case, n_threads)
        model = None
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        port = {
            "value": round(v, 4),
            "unit": "Mcell-iter/s",
            "cores": n_threads,
            "host": {"model": model, "nproc": os.cpu_count()},
            "kind": "port",
            "sample": f"one inner iteration (3 RK stages) of the full {args.config} case, "
                      f"{dt:.1f} s, FP64 C oracle with OpenMP over cubes",
        }
        ref = cpu_baseline_reference(case, B.case_params(args.config), n_threads, args.config)
        if ref is None:
            out["cpu_baseline"] = port
        else:
            rv, loop, total = ref
            out["cpu_baseline"] = {
                "value": round(rv, 4),
                "unit": "Mcell-iter/s",
                "cores": n_threads,
                "host": port["host"],
                "kind": "reference",
                "sample": f"one inner iteration (3 RK stages) of the full {args.config} case: the "
                          f"reference's own routines (oracle/build_ref.sh, -O2, OpenMP over "
                          f"{n_threads} threads), time loop {loop:.1f} s of a {total:.1f} s run "
                          f"(setup excluded)",
                "port_value": port["value"],
            }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
