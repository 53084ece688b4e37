"""Host case setup against the reference's own setup pipeline.

The fixtures were written by tests/golden/make_golden.py from the reference
sources (BCM_Grid, BCM_FWS_Interface, BCM_Initial_condition, the ghost
detection of main.cpp:694-803, BCM_Immersed_boundary, BCM_Reading_IBM,
BCM_Interface) run on the grid and STL files this generator writes.
"""
import numpy as np
import pytest

from helpers import golden, golden_ib, sha

CASES = ["tiny", "cfg1", "ml_tiny", "ml", "ml3"]


@pytest.fixture(scope="module", params=CASES)
def built(request, B):
    name = request.param
    return name, B.Case(B.case_params(name))


def test_ib_tables_match_reference(built):
    name, case = built
    ref = golden_ib(name)
    for side in ("minus", "plus"):
        gc = case.array(f"ib_{side}_gc")[1:].reshape(-1, 4)
        ip = case.array(f"ib_{side}_ip")[1:].reshape(-1, 4)
        w = case.array(f"ib_{side}_w")[1:].reshape(-1, 8)
        assert gc.shape == ref[f"{side}_gc"].shape
        assert np.array_equal(gc, ref[f"{side}_gc"])
        assert np.array_equal(ip, ref[f"{side}_ip"])
        assert np.array_equal(w, ref[f"{side}_w"])


def test_flags_and_initial_state_match_reference(built):
    name, case = built
    g = golden(name)["sha256"]
    assert sha(case.array("FWS").astype(np.int32)) == g["setup_FWS.bin"]
    assert sha(case.array("U1_")) == g["setup_U1_.bin"]
    assert sha(case.array("U1")) == g["setup_U1.bin"]
    assert sha(case.array("U1q")) == g["setup_U1q.bin"]


def test_ib_weights_interpolate_linear_fields(built):
    # Trilinear weights sum to one and reproduce the image point, to the
    # 6-decimal rounding the reference's "%f" serialisation applies.
    name, case = built
    for side in ("minus", "plus"):
        w = case.array(f"ib_{side}_w")[1:].reshape(-1, 8)
        if w.size == 0:
            continue
        assert np.all(np.abs(w.sum(axis=1) - 1.0) <= 8 * 5e-7 + 1e-12)
        assert np.all(w >= -1e-7) and np.all(w <= 1 + 1e-7)


def test_ghost_cells_are_flagged_and_interior(built):
    name, case = built
    fws = case.array("FWS")
    for side in ("minus", "plus"):
        gc = case.array(f"ib_{side}_gc")[1:].reshape(-1, 4)
        ip = case.array(f"ib_{side}_ip")[1:].reshape(-1, 4)
        assert np.all(gc[:, 1:] >= 2) and np.all(gc[:, 1:] <= 17)
        assert np.all(ip[:, 0] == gc[:, 0])           # image point in the ghost's own cube
        assert np.all(ip[:, 1:] >= 0) and np.all(ip[:, 1:] <= 18)
        f = fws[gc[:, 0], gc[:, 1], gc[:, 2], gc[:, 3]]
        assert np.all(f <= 0)


def test_case_rejects_bad_parameters(B):
    with pytest.raises(B.BCMError):
        B.Case(B.case_params("tiny", lattice=(0, 2, 2)))
    with pytest.raises(B.BCMError):
        B.Case(B.case_params("tiny", mach=-1.0))
