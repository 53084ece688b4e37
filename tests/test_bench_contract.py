"""The bench.py output contract: one JSON line with the metric, the roofline of
the dominant kernel (HBM and instruction issue) and the CPU baseline; more
than one GPU is refused before any device work."""
import json
import os
import subprocess
import sys

import pytest

from conftest import ROOT

sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_profile_figures_parse():
    import bench
    assert bench.read_traffic("k_flux_march", "cfg2") > 0
    assert bench.read_traffic("k_flux_mixed", "cfg2") > 0
    assert bench.read_valu("k_flux_march<true>") > 1000
    assert bench.read_valu("k_flux_mixed") > 1000
    v = bench.valu_roofline("fp64", 8388608, 2.0)
    per = bench.read_valu("k_flux_march<true>")
    assert v["achieved"] == pytest.approx(per * 8388608 / 2e-3 / 1e12, rel=1e-3)
    assert v["peak"] == pytest.approx(256 * 4 * 16 * 2.4e9 / 1e12, rel=1e-3)
    # At the full issue rate: cell-stages/s = issue rate / instructions, each
    # moving 161 algorithmic bytes, against 8 TB/s.
    assert v["hbm_frac_at_full_issue"] == pytest.approx(
        256 * 4 * 16 * 2.4e9 / per * 161 / 8e12, rel=1e-3)
    r = bench.roofline("fp64", 8388608, 2.0, "cfg2")
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert r["achieved"] == pytest.approx(161 * 8388608 / 2e-3 / 1e9, rel=1e-3)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0, rel=1e-3)


def test_drag_coefficient_rescales_the_reference_force():
    import math

    import bcm_cgli_amd as B
    # er[6] = Fx / (0.5 rho0 U0^2 * 2.602) (the reference's normalisation);
    # Cd = Fx / (0.5 rho0 U0^2 * pi D^2 / 4), with the sign the reference
    # prints (-er[6]).
    D = 0.25
    area = math.pi * D * D / 4
    assert B.drag_coefficient(-area / 2.602, D) == pytest.approx(1.0, rel=1e-15)
    assert B.drag_coefficient(0.0318, D) == pytest.approx(-0.0318 * 2.602 / area, rel=1e-15)
    assert B.drag_coefficient(-0.0318, 2 * D) == pytest.approx(
        B.drag_coefficient(-0.0318, D) / 4, rel=1e-15)


@pytest.mark.gpu
def test_bench_line_on_a_small_case():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--config", "cfg1", "--steps",
                        "2", "--warmup", "1", "--no-forms"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [x for x in r.stdout.splitlines() if x.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config", "roofline",
              "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 2 and d["warmup"] == 1 and d["value"] > 0
    ro = d["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in ro, k
    cb = d["cpu_baseline"]
    for k in ("value", "unit", "cores", "kind", "sample"):
        assert k in cb, k
    assert cb["kind"] in ("reference", "port") and cb["value"] > 0
