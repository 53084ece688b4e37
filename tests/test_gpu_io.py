"""GPU tests of output and restart of resident cases.

* The engine's asynchronous PLOT3D output of an FP64 case after the golden
  run is byte-identical to the files the reference's BCM_Output wrote.
* Outputs queued between steps hold the state of their own step while the
  solver runs on (two in flight), for both precisions.
* A run resumed from a restart snapshot continues bitwise as the
  uninterrupted run (FP64); the mixed path re-rounds its FP32 deviation on
  load and stays within the mixed-precision bar.
"""
import hashlib
import os

import numpy as np
import pytest

from bcm_cgli_amd.configs import STEPPING
from helpers import golden

pytestmark = pytest.mark.gpu


def file_sha(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


@pytest.mark.parametrize("name", ["tiny", "cfg1"])
def test_engine_output_matches_reference_files(B, name, tmp_path):
    case = B.Case(B.case_params(name))
    n_steps, n_inner = STEPPING[name]
    e = B.Engine(0)
    try:
        e.alloc_pool(case.ncube - 1, 1, "fp64")
        s = e.add_case(case)
        e.run(n_steps, n_inner)
        e.output(s, n_steps, str(tmp_path))
        e.output_wait()
    finally:
        e.close()
    ref = golden(name)["file_sha256"]
    for f in ("BCMdata00001.x", f"qBCMdata{n_steps:05d}.q"):
        assert file_sha(tmp_path / f) == ref[f], f


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_outputs_in_flight_hold_their_own_step(B, precision, tmp_path):
    case = B.Case(B.case_params("tiny"))
    e = B.Engine(0)
    try:
        e.alloc_pool(case.ncube - 1, 1, precision)
        s = e.add_case(case)
        states = {}
        for step in (1, 2, 3):
            e.run(1, 2)
            d = tmp_path / f"out{step}"
            d.mkdir()
            e.output(s, step, str(d))
            states[step] = e.case_state(s, "U1_")
        e.output_wait()
    finally:
        e.close()
    for step, st in states.items():
        ref_dir = tmp_path / f"ref{step}"
        ref_dir.mkdir()
        case.write_plot3d(str(ref_dir), step, st)
        for f in ("BCMdata00001.x", f"qBCMdata{step:05d}.q"):
            assert file_sha(tmp_path / f"out{step}" / f) == file_sha(ref_dir / f), (step, f)


def test_output_rejects_bad_arguments(B, tmp_path):
    case = B.Case(B.case_params("tiny"))
    e = B.Engine(0)
    try:
        e.alloc_pool(case.ncube - 1, 1, "fp64")
        s = e.add_case(case)
        with pytest.raises(B.BCMError, match="slot"):
            e.output(s + 1, 1, str(tmp_path))
        with pytest.raises(B.BCMError, match="step"):
            e.output(s, 100000, str(tmp_path))
        e.output(s, 1, str(tmp_path / "missing_dir"))
        with pytest.raises(B.BCMError, match="cannot write"):
            e.output_wait()
        e.output_wait()
    finally:
        e.close()


def _run_with_restart(B, case, precision, path, n1, n2, n_inner):
    e = B.Engine(0)
    try:
        e.alloc_pool(case.ncube - 1, 1, precision)
        s = e.add_case(case)
        e.run(n1, n_inner)
        e.save_restart(s, path, n1, E=np.arange(1.0, 6.0))
        er = e.run(n2, n_inner)[:, s, :]
        full = {n: e.case_state(s, n) for n in ("U1_", "U1", "U1q")}
    finally:
        e.close()
    e = B.Engine(0)
    try:
        e.alloc_pool(case.ncube - 1, 1, precision)
        s = e.add_case(case)
        step, E = e.load_restart(s, path)
        assert step == n1 and np.array_equal(E, np.arange(1.0, 6.0))
        er2 = e.run(n2, n_inner)[:, s, :]
        resumed = {n: e.case_state(s, n) for n in ("U1_", "U1", "U1q")}
    finally:
        e.close()
    return full, resumed, er, er2


def test_fp64_restart_resumes_bitwise(B, tmp_path):
    case = B.Case(B.case_params("tiny"))
    full, resumed, er, er2 = _run_with_restart(B, case, "fp64", str(tmp_path / "r.rst"), 2, 2, 3)
    for n in full:
        assert np.array_equal(full[n], resumed[n]), n
    assert np.array_equal(er, er2)


def test_mixed_restart_resumes_within_bar(B, tmp_path):
    case = B.Case(B.case_params("tiny"))
    full, resumed, er, er2 = _run_with_restart(B, case, "mixed", str(tmp_path / "r.rst"), 2, 2, 3)
    fluid = case.array("FWS")[1:, 2:18, 2:18, 2:18] == 1
    a = full["U1"][1:, 2:18, 2:18, 2:18, 0][fluid]
    b = resumed["U1"][1:, 2:18, 2:18, 2:18, 0][fluid]
    l2 = np.linalg.norm(a - b) / np.linalg.norm(a)
    print(f"mixed restart: density rel L2 {l2:.3e}, "
          f"bitwise {all(np.array_equal(full[n], resumed[n]) for n in full)}")
    assert l2 <= 1e-6
    assert abs(er[-1, 5] - er2[-1, 5]) <= 1e-5 * abs(er[-1, 5])
