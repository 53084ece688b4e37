"""On-disk formats and case I/O against the reference's own readers and writers.

Pins:
* the IB tables and PLOT3D files this library writes are byte-identical to
  the ones the reference's BCM_Immersed_boundary and BCM_Output wrote for the
  same case (SHA-256 in the golden fixtures; the PLOT3D solution is written
  from the oracle's final state, itself bitwise the reference's);
* a case read back from its grid and STL files equals the generated case in
  every array (the generated case is pinned to the reference's setup);
* the reference-interface drop-ins BCM_Grid and BCM_Reading_IBM fill the
  caller's arrays as the reference's readers do;
* the rank-decomposition files written for one rank read, through a
  restatement of the reference's fscanf sequence, as no cross-rank
  connections, no neighbour ranks and every cube on rank 0, and the case
  reader refuses decompositions over several ranks;
* the PLOT3D layout follows from the state by an independent numpy
  transposition; the restart snapshot round-trips bit for bit and rejects
  files that do not fit.
"""
import ctypes
import hashlib
import os

import numpy as np
import pytest

from bcm_cgli_amd.configs import STEPPING
from helpers import golden, golden_ib

CASES = ["tiny", "cfg1", "ml_tiny", "ml"]


def file_sha(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


@pytest.fixture(scope="module", params=CASES)
def built(request, B):
    name = request.param
    return name, B.Case(B.case_params(name))


def test_ibm_files_match_reference_bytes(built, tmp_path):
    name, case = built
    case.write_ibm(str(tmp_path))
    ref = golden(name)["file_sha256"]
    for f in ("IBM_minus00000.dat", "IBM_plus00000.dat"):
        assert file_sha(tmp_path / f) == ref[f], f


def test_plot3d_files_match_reference_bytes(built, O, tmp_path):
    name, case = built
    n_steps, n_inner = STEPPING[name]
    orc = O.OracleRun(case, nthreads=8)
    orc.run(n_steps, n_inner)
    case.write_plot3d(str(tmp_path), n_steps, orc.state["U1_"])
    ref = golden(name)["file_sha256"]
    for f in ("BCMdata00001.x", f"qBCMdata{n_steps:05d}.q"):
        assert file_sha(tmp_path / f) == ref[f], f


def test_plot3d_layout(B, tmp_path):
    case = B.Case(B.case_params("tiny"))
    rng = np.random.default_rng(5)
    U = rng.standard_normal((case.ncube, 20, 20, 20, 5))
    case.write_plot3d(str(tmp_path), 7, U)
    n = case.ncube - 1
    for fname, per_cube in (("BCMdata00001.x", 3 * 18 ** 3), ("qBCMdata00007.q", 4 + 5 * 18 ** 3)):
        raw = (tmp_path / fname).read_bytes()
        hdr = np.frombuffer(raw[: 4 * (1 + 3 * n)], dtype=np.int32)
        assert hdr[0] == n and np.all(hdr[1:] == 18)
        body = np.frombuffer(raw[4 * (1 + 3 * n):], dtype=np.float64)
        assert body.size == n * per_cube
    grid = np.frombuffer((tmp_path / "BCMdata00001.x").read_bytes()[4 * (1 + 3 * n):],
                         dtype=np.float64).reshape(n, 3, 18, 18, 18)
    X, Y, Z = (case.array(a)[1:, 1:19] for a in ("Xcnt", "Ycnt", "Zcnt"))
    # Block axes (a, b, c) run over cells 1..18; X follows the last, Z the first.
    assert np.array_equal(grid[:, 0], np.broadcast_to(X[:, None, None, :], grid[:, 0].shape))
    assert np.array_equal(grid[:, 1], np.broadcast_to(Y[:, None, :, None], grid[:, 1].shape))
    assert np.array_equal(grid[:, 2], np.broadcast_to(Z[:, :, None, None], grid[:, 2].shape))
    q = np.frombuffer((tmp_path / "qBCMdata00007.q").read_bytes()[4 * (1 + 3 * n):],
                      dtype=np.float64).reshape(n, 4 + 5 * 18 ** 3)
    assert np.all(q[:, :4] == 1.0)
    sol = q[:, 4:].reshape(n, 5, 18, 18, 18)
    expect = U[1:, 1:19, 1:19, 1:19, :].transpose(0, 4, 3, 2, 1)
    assert np.array_equal(sol, expect)


def test_case_from_files_equals_generated(built, tmp_path):
    name, case = built
    case.write_inputs(str(tmp_path))
    import bcm_cgli_amd as B
    back = B.Case.from_files(str(tmp_path), B.case_params(name))
    assert back.info == case.info
    assert back.phys == case.phys
    for a in ("FWS", "FWS_grid", "U1_", "U1", "U1q", "cube_size", "Xcube", "Ycube", "Zcube",
              "Xcnt", "Ycnt", "Zcnt", "csl", "adj_number", "ib_minus_gc", "ib_minus_ip",
              "ib_minus_w", "ib_plus_gc", "ib_plus_ip", "ib_plus_w", "Xbc_l", "Xbc_u", "Ybc_l",
              "Ybc_u", "Zbc_l", "Zbc_u", "nadj", "adjX_eq", "adjY_eq", "adjZ_eq",
              *(f"adj{ax}_{kind}" for ax in "XYZ"
                for kind in ("bs_plus", "sb_plus", "bs_minus", "sb_minus"))):
        assert np.array_equal(back.array(a), case.array(a)), a


def test_read_case_reports_missing_and_corrupt_files(B, tmp_path):
    p = B.case_params("tiny")
    with pytest.raises(B.BCMError, match="cannot open"):
        B.Case.from_files(str(tmp_path), p)
    case = B.Case(p)
    case.write_inputs(str(tmp_path))
    grid = tmp_path / "BCMgrid00000.dat"
    lines = grid.read_text().split("\n")
    first_adj = lines.index("adjacent cubes (4 slots) for -X +X -Y +Y -Z +Z") + 1
    lines[first_adj] = "99\t0\t0\t0"
    grid.write_text("\n".join(lines))
    with pytest.raises(B.BCMError, match="missing cube"):
        B.Case.from_files(str(tmp_path), p)
    case.write_inputs(str(tmp_path))
    stl = tmp_path / "BCM_STL_fine.stl"
    text = stl.read_text()
    stl.write_text(text[: text.index("vertex", 2000) + 8])
    with pytest.raises(B.BCMError, match="STL facet"):
        B.Case.from_files(str(tmp_path), p)


class _Scan:
    """The two fscanf directives the reference reads these files with:
    "%[^\\n]\\n" (a non-empty rest of line, then white space) and "%d\\n"."""

    def __init__(self, text):
        self.t, self.i = text, 0

    def _ws(self):
        while self.i < len(self.t) and self.t[self.i] in " \t\r\n":
            self.i += 1

    def line(self):
        end = self.t.find("\n", self.i)
        end = len(self.t) if end < 0 else end
        assert end > self.i, "empty title line"
        s, self.i = self.t[self.i:end], end
        self._ws()
        return s

    def int(self):
        self._ws()
        j = self.i + (self.t[self.i] in "+-")
        while j < len(self.t) and self.t[j].isdigit():
            j += 1
        v, self.i = int(self.t[self.i:j]), j
        self._ws()
        return v


def _read_communication(text):
    """main.cpp:262-554: per transfer kind three title lines, the neighbour
    count, then a title line before each list (rank ids, cube counts, then
    cube, direction and adjacency lists sized by the summed cube counts)."""
    sc, out = _Scan(text), []
    for n_cube_lists in (4, 5, 6):
        for _ in range(3):
            sc.line()
        ncpu = sc.int()
        sc.line()
        cpus = [sc.int() for _ in range(ncpu)]
        sc.line()
        counts = [sc.int() for _ in range(ncpu)]
        lists = []
        for _ in range(n_cube_lists):
            sc.line()
            lists.append([sc.int() for _ in range(sum(counts))])
        out.append((ncpu, cpus, counts, lists))
    assert sc.i == len(text)
    return out


def test_rank_tables_read_as_one_rank(B, tmp_path):
    p = B.case_params("tiny")
    case = B.Case(p)
    case.write_inputs(str(tmp_path))
    n = case.info["ncube"] - 1
    # main.cpp:215-221: lines up to the marker, then the connection count.
    sc = _Scan((tmp_path / "MPI_communication_table_in_turn.dat").read_text())
    while sc.line() != "Number of Cubes for connection  >>":
        pass
    assert sc.int() == 0
    for ncpu, cpus, counts, lists in _read_communication(
            (tmp_path / "Communication_irank.00000.dat").read_text()):
        assert ncpu == 0 and cpus == [] and counts == [] and all(x == [] for x in lists)
    # main.cpp:584-596 and the output offsets of BCM_Output.cpp:68-77: every
    # cube on rank 0, whose last local number is its cube count.
    sc = _Scan((tmp_path / "rank_map.dat").read_text())
    sc.line()
    rows = [(sc.int(), sc.int(), sc.int()) for _ in range(n)]
    assert sc.i == len(sc.t)
    assert [r[1] for r in rows] == [0] * n
    assert rows[-1][2] == n and sorted(r[2] for r in rows) == list(range(1, n + 1))


def test_read_case_refuses_several_ranks(B, tmp_path):
    p = B.case_params("tiny")
    case = B.Case(p)
    n = case.info["ncube"] - 1
    comm = tmp_path / "Communication_irank.00000.dat"
    rmap = tmp_path / "rank_map.dat"
    conn = tmp_path / "MPI_communication_table_in_turn.dat"

    def planted(path, edit, match):
        case.write_inputs(str(tmp_path))
        path.write_text(edit(path.read_text()))
        with pytest.raises(B.BCMError, match=match):
            B.Case.from_files(str(tmp_path), p)

    several = "several ranks"
    planted(conn, lambda t: t.replace(">>\n0\n", ">>\n12\n"), several)
    planted(conn, lambda t: t.replace(">>\n0\n", ">>\n"), "connection count")
    lines_at = lambda t, i, v: "\n".join(v if k == i else x for k, x in enumerate(t.split("\n")))
    case.write_inputs(str(tmp_path))
    zeros = [i for i, x in enumerate(comm.read_text().split("\n")) if x == "0"]
    assert len(zeros) == 3
    for at in zeros:  # neighbour count of each transfer kind
        planted(comm, lambda t: lines_at(t, at, "2"), several)
    planted(rmap, lambda t: t.replace(f"\n{n - 1}\t0\t{n}\n", f"\n{n - 1}\t1\t1\n"), several)
    planted(rmap, lambda t: t + f"{n}\t0\t{n + 1}\n", several)
    planted(rmap, lambda t: t.replace(f"{n - 1}\t0\t{n}\n", ""), f"maps {n - 1} cubes")
    planted(rmap, lambda t: t.replace("\n0\t0\t1\n", "\n0\t0\t0\n"), "local cube number")
    planted(rmap, lambda t: t + "x\n", "row unreadable")
    # Absent decomposition files are accepted; the written ones read back.
    case.write_inputs(str(tmp_path))
    B.Case.from_files(str(tmp_path), p)
    for f in (comm, rmap, conn):
        f.unlink()
    assert B.Case.from_files(str(tmp_path), p).info == case.info


def _dropin(B):
    lib = B.load_library()
    P = ctypes.c_void_p
    lib.BCM_Grid.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int] + [P] * 11
    lib.BCM_Grid.restype = None
    lib.BCM_Reading_IBM.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int, P, P,
                                    P]
    lib.BCM_Reading_IBM.restype = None
    return lib


def test_bcm_grid_dropin_fills_reference_arrays(B, tmp_path, monkeypatch):
    case = B.Case(B.case_params("cfg1"))
    case.write_inputs(str(tmp_path))
    monkeypatch.chdir(tmp_path)
    lib = _dropin(B)
    nc, nw = case.ncube, case.info["n_wallcube"]
    arrs = dict(cube_size=np.full(nc, -7.0), csl=np.full(nc, -7, np.int32),
                Xcube=np.full(nc, -7.0), Ycube=np.full(nc, -7.0), Zcube=np.full(nc, -7.0),
                Xcnt=np.full((nc, 20), -7.0), Ycnt=np.full((nc, 20), -7.0),
                Zcnt=np.full((nc, 20), -7.0), FWS=np.full((nc, 20, 20, 20), -7, np.int32),
                adj_number=np.full((nc, 5, 7), -7, np.int32), wallcube=np.full(nc, -7, np.int32))
    order = ("cube_size", "csl", "Xcube", "Ycube", "Zcube", "Xcnt", "Ycnt", "Zcnt", "FWS",
             "adj_number", "wallcube")
    lib.BCM_Grid(0, nc, nw, *(arrs[k].ctypes.data for k in order))
    for k in ("cube_size", "csl", "Xcube", "Ycube", "Zcube", "Xcnt", "Ycnt", "Zcnt"):
        assert np.array_equal(arrs[k][1:], case.array(k)[1:]), k
        assert np.all(arrs[k][0] == -7), k
    assert np.array_equal(arrs["FWS"][1:], case.array("FWS_grid")[1:])
    adj = case.array("adj_number")
    assert np.array_equal(arrs["adj_number"][1:, 1:5, 1:7], adj[1:, 1:5, 1:7])
    assert np.all(arrs["adj_number"][:, 0, :] == -7) and np.all(arrs["adj_number"][:, :, 0] == -7)
    # Wall cubes: every cube holding a solid interior cell, in cube order.
    fg = case.array("FWS_grid")[1:, 2:18, 2:18, 2:18]
    walls = [c + 1 for c in range(nc - 1) if (fg[c] == -1).any()]
    assert nw - 1 == len(walls)
    assert arrs["wallcube"][1:nw].tolist() == walls


def test_bcm_reading_ibm_dropin_reads_reference_tables(B, tmp_path, monkeypatch):
    name = "tiny"
    case = B.Case(B.case_params(name))
    case.write_ibm(str(tmp_path))
    monkeypatch.chdir(tmp_path)
    lib = _dropin(B)
    ref = golden_ib(name)
    for side, sw in (("minus", -1), ("plus", 1)):
        n = ref[f"{side}_gc"].shape[0]
        gc = np.zeros(4 * n + 1, np.int32)
        ip = np.zeros(4 * n + 1, np.int32)
        w = np.zeros(8 * n + 1)
        nn = ctypes.c_int(n)
        lib.BCM_Reading_IBM(0, ctypes.byref(nn), sw, w.ctypes.data, gc.ctypes.data, ip.ctypes.data)
        assert np.array_equal(gc[1:].reshape(-1, 4), ref[f"{side}_gc"])
        assert np.array_equal(ip[1:].reshape(-1, 4), ref[f"{side}_ip"])
        assert np.array_equal(w[1:].reshape(-1, 8), ref[f"{side}_w"])


def test_restart_round_trip(B, tmp_path):
    case = B.Case(B.case_params("tiny"))
    rng = np.random.default_rng(11)
    st = [rng.standard_normal((case.ncube, 20, 20, 20, 5)) for _ in range(3)]
    E = np.array([1.5, 2.5, 3.5, 4.5, 5.5])
    path = str(tmp_path / "case.rst")
    B.write_restart(path, 1234, *st, E=E)
    step, e, *back = B.read_restart(path, case.ncube)
    assert step == 1234 and np.array_equal(e, E)
    for a, b in zip(st, back):
        assert np.array_equal(a, b)
    with pytest.raises(B.BCMError, match="cube slots"):
        B.read_restart(path, case.ncube + 1)
    raw = open(path, "rb").read()
    with open(path, "wb") as f:
        f.write(raw[:-8])
    with pytest.raises(B.BCMError, match="truncated"):
        B.read_restart(path, case.ncube)
    with open(path, "wb") as f:
        f.write(b"NOTARST!" + raw[8:])
    with pytest.raises(B.BCMError, match="not a restart"):
        B.read_restart(path, case.ncube)
