"""GPU parity: the HIP path through the C ABI against the oracle and the
reference fixtures.  Bar: bitwise equality for every state array; the norms
er[1..7] are block-reduced on the device in a different summation order than
the reference's sequential sums, so they are compared at 1e-12 relative.
"""
import ctypes
import hashlib
import os

import numpy as np
import pytest

from bcm_cgli_amd.configs import STEPPING
from helpers import describe_mismatch, golden, sha

pytestmark = pytest.mark.gpu

STATE = ("U1_", "U1", "U1q", "U1p1", "U1p2", "Rku1", "Residual1")
ER_RTOL = 1e-12


def er_close(a, b):
    """er[1..5] (change norms) within 1e-12 of each column's magnitude; the two
    force sums er[6..7] within 1e-12 of the force magnitude max|er[6..7]|:
    their terms cancel (er[7] is a near-zero side force), so the summation
    order changes them by round-off of the terms, not of the total."""
    a = np.atleast_2d(np.asarray(a))
    b = np.atleast_2d(np.asarray(b))
    scale = np.maximum(np.abs(b).max(axis=0), 1e-300)
    scale[5:7] = max(np.abs(b[..., 5:7]).max(), 1e-300)
    ok = np.all(np.abs(a - b) <= ER_RTOL * scale)
    if not ok:
        print("er max deviation / scale per column:", (np.abs(a - b) / scale).max(axis=0))
    return ok


@pytest.fixture(scope="module")
def engine(B):
    e = B.Engine(0)
    yield e
    e.close()


def compare(eng, orc, names):
    for n in names:
        a = eng.get_state(n)
        b = orc.state[n]
        if not np.array_equal(a.view(np.int64), b.view(np.int64)):
            raise AssertionError(f"{n}: {describe_mismatch(a, b)}")


@pytest.mark.parametrize("name", ["tiny", "ml_tiny"])
def test_each_kernel_bitwise(B, O, engine, name):
    case = B.Case(B.case_params(name))
    engine.load_case(case)
    orc = O.OracleRun(case, nthreads=8)
    compare(engine, orc, ("U1_", "U1", "U1q"))
    for RK in (1, 2, 3):
        engine.part("bc", RK); orc.bc(); compare(engine, orc, ("U1_",))
        engine.part("ghost", RK); orc.ghost(); compare(engine, orc, ("U1_",))
        # pre-flux exchange as the engine runs it: IB ghost copies + edge passes
        engine.part("halo_refresh", RK); orc.interface(); compare(engine, orc, ("U1_",))
        engine.part("flux_update", RK); orc.flux_update(RK)
        compare(engine, orc, ("U1_", "Rku1", "Residual1", "U1p1", "U1p2"))
        engine.part("interface", RK); orc.interface(); compare(engine, orc, ("U1_",))
    assert er_close(engine.er()[1:], orc.er[1:])
    engine.part("shift"); orc.shift(); compare(engine, orc, STATE)


@pytest.mark.parametrize("name", ["tiny", "cfg1", "ml_tiny", "ml", "ml3"])
def test_run_matches_reference_fixture(B, engine, name):
    g = golden(name)
    case = B.Case(B.case_params(name))
    engine.load_case(case)
    n_steps, n_inner = STEPPING[name]
    hist = engine.run(n_steps, n_inner)
    for n in STATE:
        assert sha(engine.get_state(n)) == g["sha256"][f"final_{n}.bin"], n
    assert er_close(hist, g["er"])


FLUX_CHOICES = [(0, 1, 0), (1, 1, 1), (1, 1, 2), (1, 1, 0), (1, 2, 0), (1, 4, 0), (1, 16, 0)]


@pytest.mark.parametrize("name", ["tiny", "ml_tiny"])
@pytest.mark.parametrize("variant,nseg,fuse", FLUX_CHOICES)
def test_flux_variants_bitwise(B, O, engine, variant, nseg, fuse, name):
    # Every flux kernel choice the autotuner may pick gives the oracle's bits,
    # including cubes where neighbouring cells choose different face orders,
    # and the RK update fused behind the marching flux; on one-size and
    # two-level grids.
    case = B.Case(B.case_params(name))
    engine.load_case(case)
    engine.set_flux(variant, nseg, fuse)
    orc = O.OracleRun(case, nthreads=8)
    h = engine.run(1, 2)
    ho = orc.run(1, 2)
    compare(engine, orc, STATE)
    assert er_close(h, ho)


@pytest.mark.parametrize("name", ["tiny", "cfg1", "ml_tiny", "ml", "ml3"])
@pytest.mark.parametrize("use_graph", [True, False])
def test_push_stages_match_reference_fixture(B, engine, name, use_graph):
    # Stages whose flux kernel writes the face halos into the other copy of
    # the iterate (odd flux count per step for tiny: copy-back path).
    g = golden(name)
    case = B.Case(B.case_params(name))
    engine.load_case(case)
    engine.set_flux(1, 1, 2)
    n_steps, n_inner = STEPPING[name]
    hist = engine.run(n_steps, n_inner, use_graph=use_graph)
    for n in STATE:
        assert sha(engine.get_state(n)) == g["sha256"][f"final_{n}.bin"], n
    assert er_close(hist, g["er"])


def test_autotune_choice_is_bitwise(B, O, engine):
    case = B.Case(B.case_params("cfg1"))
    engine.load_case(case)
    choice = engine.autotune()
    assert choice["variant"] in (0, 1)
    # tuning restores the state it timed on
    assert sha(engine.get_state("U1_")) == sha(case.array("U1_"))
    orc = O.OracleRun(case, nthreads=8)
    engine.run(1, 1)
    orc.run(1, 1)
    compare(engine, orc, STATE)


def test_full_exchange_after_external_state(B, O, engine):
    # After the inner iterate is replaced from outside, the first pre-flux
    # exchange must be the full one; the oracle always runs the full one.
    case = B.Case(B.case_params("tiny"))
    engine.load_case(case)
    rng = np.random.default_rng(5)
    U = case.array("U1_").copy()
    U[1:, 2:18, 2:18, 2:18, 0] *= 1 + 1e-3 * rng.uniform(-1, 1, size=U[1:, 2:18, 2:18, 2:18, 0].shape)
    engine.set_state("U1_", U)
    orc = O.OracleRun(case, nthreads=8)
    orc.state["U1_"][...] = U
    engine.run(1, 2)
    orc.run(1, 2)
    compare(engine, orc, STATE)


def test_graph_replays_match_eager(B, engine):
    # Several replays of the captured inner iteration, with eager launches and
    # host synchronisation in between, against an eager-only run.
    case = B.Case(B.case_params("tiny"))
    eager = B.Engine(0)
    eager.load_case(case)
    engine.load_case(case)
    for rep in range(3):
        h_graph = engine.run(1, 2, use_graph=True)
        engine.part("bc", 1)  # eager work between replays; repeated below
        h_eager = eager.run(1, 2, use_graph=False)
        eager.part("bc", 1)
        for n in STATE:
            a, b = engine.get_state(n), eager.get_state(n)
            assert np.array_equal(a.view(np.int64), b.view(np.int64)), (rep, n)
        assert np.array_equal(h_graph.view(np.int64), h_eager.view(np.int64))
    eager.close()


EDGE_CASES = {
    # one cube: every face is a domain boundary, no interfaces
    "single_cube": dict(lattice=(1, 1, 1), cube_size=0.5, sphere_center=(0.25, 0.25, 0.25)),
    # no body at all: no ghost cells, empty IB lists
    "no_body": dict(lattice=(2, 1, 1), sphere_diameter=0.0, sphere_center=(0.1, 0.1, 0.1)),
    # non-cubic lattice with the sphere cut by the z-min symmetry plane
    "ragged_cut": dict(lattice=(3, 2, 1), sphere_center=(0.4, 0.25, 0.05)),
    # sphere inside a single cube of a 2x2x2 lattice, coarse surface
    "inside_one_cube": dict(sphere_center=(0.125, 0.375, 0.125), sphere_diameter=0.15, ico_level=2),
}


@pytest.mark.parametrize("label", sorted(EDGE_CASES))
def test_edge_cases_bitwise(B, O, engine, label):
    case = B.Case(B.case_params("tiny", **EDGE_CASES[label]))
    engine.load_case(case)
    orc = O.OracleRun(case, nthreads=8)
    h = engine.run(2, 2)
    ho = orc.run(2, 2)
    compare(engine, orc, STATE)
    assert er_close(h, ho)


def test_cfg2_one_inner_iteration_bitwise(B, O, engine):
    # Full-size case (16x16x8 cubes, 8.4M cells): one inner iteration against
    # the multi-threaded oracle, plus a size-independent property: a second
    # identical run from the same start is bit-identical (determinism).
    case = B.Case(B.case_params("cfg2"))
    engine.load_case(case)
    h1 = engine.run(1, 1)
    orc = O.OracleRun(case, nthreads=16)
    ho = orc.run(1, 1)
    compare(engine, orc, ("U1_", "U1", "U1q", "Rku1", "Residual1"))
    assert er_close(h1, ho)
    engine.load_case(case)
    h2 = engine.run(1, 1)
    assert np.array_equal(h1.view(np.int64), h2.view(np.int64))
    assert sha(engine.get_state("U1_")) == sha(orc.state["U1_"])


def test_cfg2_one_physical_step_bitwise(B, O, engine):
    # Full-size case, one physical step of 10 inner iterations as the bench
    # runs it (one captured graph: 30 pushed stages, the run-end edge pass and
    # the physical-time shift) against the oracle on the host cores: the three
    # time levels bitwise, the norms of every inner iteration at 1e-12.
    case = B.Case(B.case_params("cfg2"))
    engine.load_case(case)
    h = engine.run(1, 10)
    orc = O.OracleRun(case, nthreads=16)
    ho = orc.run(1, 10)
    compare(engine, orc, ("U1_", "U1", "U1q"))
    assert er_close(h, ho)


def test_cfg2_three_physical_steps_bitwise(B, O, engine):
    # Full-size case over three physical steps of 10 inner iterations (three
    # graph replays, 90 pushed stages, the physical-time shift between steps,
    # so the dual-time source of steps 2 and 3 comes from shifted levels)
    # against the oracle on the host cores: every array the reference keeps
    # bitwise, including the auxiliary arrays of the last inner iteration, and
    # the norms of every inner iteration at 1e-12.
    case = B.Case(B.case_params("cfg2"))
    engine.load_case(case)
    h = engine.run(3, 10)
    orc = O.OracleRun(case, nthreads=16)
    ho = orc.run(3, 10)
    compare(engine, orc, STATE)
    assert er_close(h, ho)


def _dropin_run(B, case, n_steps, n_inner, output_dir=None):
    """Runs the reference main-loop call sequence (main.cpp:1146-1357) through
    the drop-in BCM_* entry points on host arrays bound to an engine; with
    output_dir, BCM_Output writes the last step there as main.cpp:1346-1351
    does (from the bound device state)."""
    lib = B.load_library()
    c_int, c_double = ctypes.c_int, ctypes.c_double
    P = ctypes.c_void_p
    nc = case.ncube
    arrays = {n: np.zeros((nc, 20, 20, 20, 5)) for n in STATE}
    for n in ("U1_", "U1", "U1q"):
        arrays[n][...] = case.array(n)
    fws = np.ascontiguousarray(case.array("FWS"), dtype=np.int32)
    cs = np.ascontiguousarray(case.array("cube_size"))
    csl = np.ascontiguousarray(case.array("csl"))
    adj = np.ascontiguousarray(case.array("adj_number"))
    ph = case.phys
    phys7 = np.array([ph["rho0"], ph["U0"], ph["Uin"], ph["V0"], ph["W0"], ph["P0"], ph["mu_L"]])
    h = ctypes.c_void_p()
    assert lib.bcm_engine_create(0, ctypes.byref(h)) == 0
    lib.bcm_bind.argtypes = [P, c_int, c_int] + [P] * 10
    rc = lib.bcm_bind(h, nc, nc - 1, *[arrays[n].ctypes.data for n in STATE], fws.ctypes.data,
                      cs.ctypes.data, phys7.ctypes.data)
    assert rc == 0, lib.bcm_last_error()
    # Interface tables through the drop-in table builder.
    lists = [np.zeros(nc + 1, dtype=np.int32) for _ in range(15)]
    counts = [c_int() for _ in range(15)]
    lib.BCM_Interface_table.argtypes = [c_int] * 3 + [P] * 17 + [P] * 15
    lib.BCM_Interface_table(0, nc, 1, *[ctypes.byref(c) for c in counts], csl.ctypes.data,
                            adj.ctypes.data, *[l.ctypes.data for l in lists])
    bcl = [np.ascontiguousarray(case.array(n)) for n in ("Xbc_l", "Xbc_u", "Ybc_l", "Ybc_u", "Zbc_l", "Zbc_u")]
    ib = {s: [np.ascontiguousarray(case.array(f"ib_{s}_{k}")) for k in ("w", "gc", "ip")] for s in ("minus", "plus")}
    nbc = {s: c_int(case.info[f"nbc_{s}"]) for s in ("minus", "plus")}
    dummy = np.zeros(64, dtype=np.int32)
    dd = np.zeros(64)
    er = np.zeros(8)
    lib.BCM_X_boundary_condition.argtypes = [c_int, c_int, c_int, P, P, P]
    lib.BCM_Y_boundary_condition.argtypes = [c_int, c_int, P, P, P]
    lib.BCM_Z_boundary_condition.argtypes = [c_int, c_int, P, P, P]
    for fn in (lib.BCM_Ghostcell_minus, lib.BCM_Ghostcell_plus):
        fn.argtypes = [c_int, P, P, P, P, P, P]
    lib.BCM_Flux_XYZ_Viscous_Runge_kutta.argtypes = [c_int, c_int, c_int, c_double, c_double, c_double] + [P] * 11
    lib.BCM_Physical_time_shift.argtypes = [P, P, P]
    U = arrays["U1_"].ctypes.data
    di, dd_ = dummy.ctypes.data, dd.ctypes.data
    # rank_map .. Ncube_Ncpu_bs (13), then per eq/sb/bs: 2 int*, 4 double*, 2 int*,
    # then ist_* and the three adjacency helpers (6).
    mpi_ptrs = [di] * 13 + ([di] * 2 + [dd_] * 4 + [di] * 2) * 3 + [di] * 6
    assert len(mpi_ptrs) == 43

    def interface():
        args = [0, nc, 1, 0, 0, 0, 0, 0, 0] + [c.value for c in counts]
        tail = [csl.ctypes.data, adj.ctypes.data] + [l.ctypes.data for l in lists] + [U]
        fn = lib.BCM_Interface
        fn.argtypes = [c_int] * len(args) + [P] * (len(mpi_ptrs) + len(tail))
        fn(*args, *mpi_ptrs, *tail)

    hist = []
    for step in range(n_steps):
        for it in range(n_inner):
            for RK in (1, 2, 3):
                lib.BCM_X_boundary_condition(0, bcl[0].size - 1, bcl[1].size - 1, bcl[0].ctypes.data, bcl[1].ctypes.data, U)
                lib.BCM_Y_boundary_condition(bcl[2].size - 1, bcl[3].size - 1, bcl[2].ctypes.data, bcl[3].ctypes.data, U)
                lib.BCM_Z_boundary_condition(bcl[4].size - 1, bcl[5].size - 1, bcl[4].ctypes.data, bcl[5].ctypes.data, U)
                lib.BCM_Ghostcell_minus(0, ctypes.byref(nbc["minus"]), ib["minus"][0].ctypes.data,
                                        ib["minus"][1].ctypes.data, ib["minus"][2].ctypes.data, fws.ctypes.data, U)
                lib.BCM_Ghostcell_plus(0, ctypes.byref(nbc["plus"]), ib["plus"][0].ctypes.data,
                                       ib["plus"][1].ctypes.data, ib["plus"][2].ctypes.data, fws.ctypes.data, U)
                interface()
                lib.BCM_Flux_XYZ_Viscous_Runge_kutta(0, nc, RK, 1e-4, 1e-4 / 200.0, 0.1, fws.ctypes.data,
                                                     csl.ctypes.data, cs.ctypes.data,
                                                     *[arrays[n].ctypes.data for n in STATE], er.ctypes.data)
                interface()
            hist.append(er[1:8].copy())
        lib.BCM_Physical_time_shift(U, arrays["U1"].ctypes.data, arrays["U1q"].ctypes.data)
    if output_dir is not None:
        xyz = [np.ascontiguousarray(case.array(n)) for n in ("Xcnt", "Ycnt", "Zcnt")]
        lib.BCM_Output.argtypes = [c_int, c_int, c_int] + [P] * 7
        lib.BCM_Output.restype = None
        cwd = os.getcwd()
        os.chdir(output_dir)
        try:
            lib.BCM_Output(0, nc, n_steps, di, U, arrays["U1q"].ctypes.data, cs.ctypes.data,
                           *(a.ctypes.data for a in xyz))
        finally:
            os.chdir(cwd)
    assert lib.bcm_sync_to_host(h) == 0
    lib.bcm_engine_destroy(h)
    return arrays, np.array(hist)


def test_dropin_main_loop_two_level_grid(B):
    # The drop-in BCM_Interface with two-level lists (injection into small
    # cubes, restriction into big ones) against the reference fixture.
    g = golden("ml_tiny")
    case = B.Case(B.case_params("ml_tiny"))
    arrays, hist = _dropin_run(B, case, *STEPPING["ml_tiny"])
    for n in STATE:
        assert sha(arrays[n]) == g["sha256"][f"final_{n}.bin"], n
    assert er_close(hist, g["er"])


def test_dropin_main_loop_matches_reference_fixture(B, tmp_path):
    g = golden("tiny")
    case = B.Case(B.case_params("tiny"))
    n_steps, n_inner = STEPPING["tiny"]
    arrays, hist = _dropin_run(B, case, n_steps, n_inner, output_dir=str(tmp_path))
    for n in STATE:
        assert sha(arrays[n]) == g["sha256"][f"final_{n}.bin"], n
    assert er_close(hist, g["er"])
    for f in ("BCMdata00001.x", f"qBCMdata{n_steps:05d}.q"):
        with open(tmp_path / f, "rb") as fh:
            assert hashlib.sha256(fh.read()).hexdigest() == g["file_sha256"][f], f


@pytest.mark.parametrize("limit", [0, 40])
def test_ib_fill_memory_schedule_bitwise(B, O, engine, limit):
    # A cube whose IB fill needs more value slots than fit in LDS takes the
    # in-place schedule on memory (levels ordered by writes after reads as
    # well).  Lowering the slot limit puts every IB cube (limit 0), or the
    # larger ones (limit 40), on that schedule: the run is still bitwise the
    # oracle, and the mixed path gives the same bits as with its LDS fill.
    case = B.Case(B.case_params("cfg1"))
    engine.load_case(case)
    levels_lds = engine.ghost_stats()["levels"]
    try:
        engine.set_ghost_lds_limit(limit)
        engine.load_case(case)
        stats = engine.ghost_stats()
        if limit == 0:
            assert stats["levels"] > levels_lds
        orc = O.OracleRun(case, nthreads=16)
        h = engine.run(2, 3)
        ho = orc.run(2, 3)
        compare(engine, orc, STATE)
        assert er_close(h, ho)
        mixed = []
        for lim in (1536, limit):
            engine.set_ghost_lds_limit(lim)
            engine.alloc_pool(case.ncube - 1, 1, "mixed")
            s = engine.add_case(case)
            er = engine.run(2, 3)[:, s, :]
            mixed.append((engine.case_state(s, "U1_"), er))
        assert np.array_equal(mixed[0][0].view(np.int64), mixed[1][0].view(np.int64))
        assert np.array_equal(mixed[0][1].view(np.int64), mixed[1][1].view(np.int64))
    finally:
        engine.set_ghost_lds_limit(1536)
