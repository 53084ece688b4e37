"""ctypes binding of the FP64 CPU oracle (oracle/bcm_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker -- never as the measured or
shipped path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


class OrcPhys(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("rho0", "U0", "Uin", "V0", "W0", "P0", "T0", "mu_L")]


IP = ctypes.POINTER(ctypes.c_int)
DP = ctypes.POINTER(ctypes.c_double)


class OrcCase(ctypes.Structure):
    _fields_ = [
        ("ncube", ctypes.c_int),
        ("mpi_ncube", ctypes.c_int),
        ("FWS", IP),
        ("cube_size", DP),
        ("adj_number", IP),
        ("neq", ctypes.c_int * 3),
        ("eq", IP * 3),
        ("nbs", ctypes.c_int * 6),
        ("bs", IP * 6),
        ("nbc", ctypes.c_int * 6),
        ("bc", IP * 6),
        ("nib", ctypes.c_int * 2),
        ("ib_weight", DP * 2),
        ("ib_gc", IP * 2),
        ("ib_ip", IP * 2),
        ("phys", OrcPhys),
    ]


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing; run __graft_entry__.build()")
        lib = ctypes.CDLL(LIB_PATH)
        vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        lib.orc_bc_x.argtypes = [ci, ci, vp, vp, vp, ctypes.POINTER(OrcPhys)]
        lib.orc_bc_y.argtypes = [ci, ci, vp, vp, vp]
        lib.orc_bc_z.argtypes = [ci, ci, vp, vp, vp]
        lib.orc_ghostcell.argtypes = [ci, vp, vp, vp, vp, ctypes.POINTER(OrcPhys)]
        lib.orc_interface.argtypes = [ci, vp, vp, vp, vp, vp, vp]
        lib.orc_flux_rk.argtypes = [ci, ci, ci, cd, cd, cd, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                    ctypes.POINTER(OrcPhys), ci]
        lib.orc_shift.argtypes = [ci, vp, vp, vp]
        lib.orc_stage.argtypes = [ctypes.POINTER(OrcCase), ci, cd, cd, cd, vp, vp, vp, vp, vp, vp,
                                  vp, vp, ci]
        lib.orc_run.argtypes = [ctypes.POINTER(OrcCase), ci, ci, cd, cd, cd, vp, vp, vp, vp, vp, vp,
                                vp, vp, ci]
        lib.orc_run_converge.argtypes = [ctypes.POINTER(OrcCase), ci, ci, cd, vp, vp, cd, cd, cd, vp,
                                         vp, vp, vp, vp, vp, vp, vp, ci]
        _lib = lib
    return _lib


def _iptr(a):
    return a.ctypes.data_as(IP)


def _dptr(a):
    return a.ctypes.data_as(DP)


class OracleRun:
    """Oracle state for one host case: copies of the case state plus scratch arrays."""

    def __init__(self, case, nthreads=None):
        lib = load()
        self.lib = lib
        self.case = case
        self.nthreads = nthreads or os.cpu_count() or 1
        self.keep = {}
        nc = case.ncube

        def keep(name, arr):
            arr = np.ascontiguousarray(arr)
            self.keep[name] = arr
            return arr

        oc = OrcCase()
        oc.ncube = nc
        oc.mpi_ncube = nc - 1
        oc.FWS = _iptr(keep("FWS", case.array("FWS").copy()))
        oc.cube_size = _dptr(keep("cube_size", case.array("cube_size").copy()))
        oc.adj_number = _iptr(keep("adj", case.array("adj_number").copy()))
        nadj = case.array("nadj").copy()
        for d, n in enumerate(("adjX_eq", "adjY_eq", "adjZ_eq")):
            oc.neq[d] = int(nadj[d])
            oc.eq[d] = _iptr(keep(n, case.array(n).copy()))
        # Big cubes with four small neighbours on their + / - side, per axis
        # (list indices 3 + 4d and 5 + 4d of the 15 interface lists).
        for d, ax in enumerate("XYZ"):
            for s, (n, li) in enumerate(((f"adj{ax}_bs_plus", 3 + 4 * d),
                                         (f"adj{ax}_bs_minus", 5 + 4 * d))):
                oc.nbs[2 * d + s] = int(nadj[li])
                oc.bs[2 * d + s] = _iptr(keep(n, case.array(n).copy()))
        for s, n in enumerate(("Xbc_l", "Xbc_u", "Ybc_l", "Ybc_u", "Zbc_l", "Zbc_u")):
            a = keep(n, case.array(n).copy())
            oc.nbc[s] = a.size - 1
            oc.bc[s] = _iptr(a)
        for s, n in enumerate(("minus", "plus")):
            oc.nib[s] = case.info["nbc_" + n]
            oc.ib_weight[s] = _dptr(keep(n + "_w", case.array(f"ib_{n}_w").copy()))
            oc.ib_gc[s] = _iptr(keep(n + "_gc", case.array(f"ib_{n}_gc").copy()))
            oc.ib_ip[s] = _iptr(keep(n + "_ip", case.array(f"ib_{n}_ip").copy()))
        ph = OrcPhys(*[case.phys[k] for k in ("rho0", "U0", "Uin", "V0", "W0", "P0", "T0", "mu_L")])
        oc.phys = ph
        self.oc = oc
        self.phys = ph
        shape = (nc, 20, 20, 20, 5)
        self.state = {
            "U1_": case.array("U1_").copy(),
            "U1": case.array("U1").copy(),
            "U1q": case.array("U1q").copy(),
            "U1p1": np.zeros(shape),
            "U1p2": np.zeros(shape),
            "Rku1": np.zeros(shape),
            "Residual1": np.zeros(shape),
        }
        self.er = np.zeros(8)
        self.deltaT, self.deltaTau, self.e = 1e-4, 1e-4 / 200.0, 0.1

    def _s(self, n):
        return self.state[n].ctypes.data

    def bc(self):
        oc, s = self.oc, self._s("U1_")
        self.lib.orc_bc_x(oc.nbc[0], oc.nbc[1], oc.bc[0], oc.bc[1], s, ctypes.byref(self.phys))
        self.lib.orc_bc_y(oc.nbc[2], oc.nbc[3], oc.bc[2], oc.bc[3], s)
        self.lib.orc_bc_z(oc.nbc[4], oc.nbc[5], oc.bc[4], oc.bc[5], s)

    def ghost(self):
        oc = self.oc
        for s in range(2):
            self.lib.orc_ghostcell(oc.nib[s], oc.ib_weight[s], oc.ib_gc[s], oc.ib_ip[s],
                                   self._s("U1_"), ctypes.byref(self.phys))

    def interface(self):
        oc = self.oc
        neq = (ctypes.c_int * 3)(*oc.neq)
        eq = (IP * 3)(*oc.eq)
        nbs = (ctypes.c_int * 6)(*oc.nbs)
        bs = (IP * 6)(*oc.bs)
        self.lib.orc_interface(oc.ncube, neq, eq, nbs, bs, oc.adj_number, self._s("U1_"))

    def flux_update(self, RK):
        oc = self.oc
        self.lib.orc_flux_rk(oc.ncube, oc.mpi_ncube, RK, self.deltaT, self.deltaTau, self.e, oc.FWS,
                             oc.cube_size, self._s("U1_"), self._s("U1"), self._s("U1q"),
                             self._s("U1p1"), self._s("U1p2"), self._s("Rku1"),
                             self._s("Residual1"), self.er.ctypes.data, ctypes.byref(self.phys),
                             self.nthreads)

    def shift(self):
        self.lib.orc_shift(self.oc.ncube, self._s("U1"), self._s("U1q"), self._s("U1_"))

    def stage(self, RK):
        self.lib.orc_stage(ctypes.byref(self.oc), RK, self.deltaT, self.deltaTau, self.e,
                           self._s("U1_"), self._s("U1"), self._s("U1q"), self._s("U1p1"),
                           self._s("U1p2"), self._s("Rku1"), self._s("Residual1"),
                           self.er.ctypes.data, self.nthreads)

    def run(self, n_steps, n_inner):
        hist = np.zeros((n_steps * n_inner, 7))
        self.lib.orc_run(ctypes.byref(self.oc), n_steps, n_inner, self.deltaT, self.deltaTau,
                         self.e, self._s("U1_"), self._s("U1"), self._s("U1q"), self._s("U1p1"),
                         self._s("U1p2"), self._s("Rku1"), self._s("Residual1"),
                         hist.ctypes.data, self.nthreads)
        return hist

    def run_converge(self, n_steps, max_inner, tol, E=None):
        """Steps under the reference's inner-iteration control; returns
        (iterations per step, er[1..7] per inner iteration with rows after
        each step's break left zero, running maxima E1..E5)."""
        assert max_inner >= 1
        E = np.zeros(5) if E is None else np.array(E, dtype=np.float64)
        iters = np.zeros(n_steps, dtype=np.int32)
        hist = np.zeros((n_steps * max_inner, 7))
        self.lib.orc_run_converge(ctypes.byref(self.oc), n_steps, max_inner, tol, E.ctypes.data,
                                  iters.ctypes.data, self.deltaT, self.deltaTau, self.e,
                                  self._s("U1_"), self._s("U1"), self._s("U1q"), self._s("U1p1"),
                                  self._s("U1p2"), self._s("Rku1"), self._s("Residual1"),
                                  hist.ctypes.data, self.nthreads)
        return iters, hist, E
