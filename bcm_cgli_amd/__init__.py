"""MI355X-native per-iteration solver path of the BCM_CGLI Building-Cube solver.

Python binding (ctypes) of ``libbcm_cgli.so``, whose C ABI is declared in
``include/bcm_cgli.h``.  The compute path is HIP on the GPU; this module only
moves arguments and numpy views across the boundary and raises if the native
library is missing.
"""
import ctypes
import os

import numpy as np

from .configs import CONFIGS, case_params  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbcm_cgli.so")

STATE_NAMES = ("U1_", "U1", "U1q", "U1p1", "U1p2", "Rku1", "Residual1")
PARTS = {"bc": 0, "ghost": 1, "interface": 2, "flux": 3, "update": 4, "shift": 5, "stage": 6,
         "flux_update": 7, "halo_refresh": 8, "flux_push": 9}


class BCMError(RuntimeError):
    pass


class CaseParamsC(ctypes.Structure):
    _fields_ = [
        ("lattice", ctypes.c_int * 3),
        ("cube_size", ctypes.c_double),
        ("origin", ctypes.c_double * 3),
        ("sphere_center", ctypes.c_double * 3),
        ("sphere_diameter", ctypes.c_double),
        ("ico_level", ctypes.c_int),
        ("mach", ctypes.c_double),
        ("reynolds", ctypes.c_double),
        ("seed", ctypes.c_ulonglong),
        ("noise_amp", ctypes.c_double),
        ("refine_lo", ctypes.c_int * 3),
        ("refine_hi", ctypes.c_int * 3),
        ("refine2_lo", ctypes.c_int * 3),
        ("refine2_hi", ctypes.c_int * 3),
    ]


_lib = None


def load_library(path=LIB_PATH):
    """Loads the native library; raises BCMError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise BCMError(f"native library {path} is missing; run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    c_int, c_double, c_void_p, c_char_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_char_p
    P = ctypes.POINTER
    lib.bcm_last_error.restype = c_char_p
    lib.bcm_case_build.argtypes = [P(CaseParamsC), P(c_void_p)]
    lib.bcm_case_free.argtypes = [c_void_p]
    lib.bcm_case_free.restype = None
    lib.bcm_case_info.argtypes = [c_void_p, P(c_int), P(c_double)]
    lib.bcm_case_array.argtypes = [c_void_p, c_char_p, P(ctypes.c_longlong)]
    lib.bcm_case_array.restype = c_void_p
    lib.bcm_case_write_inputs.argtypes = [c_void_p, c_char_p]
    lib.bcm_engine_create.argtypes = [c_int, P(c_void_p)]
    lib.bcm_engine_destroy.argtypes = [c_void_p]
    lib.bcm_engine_destroy.restype = None
    lib.bcm_engine_load_case.argtypes = [c_void_p, c_void_p]
    lib.bcm_engine_set_time.argtypes = [c_void_p, c_double, c_double, c_double]
    lib.bcm_engine_run.argtypes = [c_void_p, c_int, c_int, c_void_p, c_int]
    lib.bcm_engine_run_converge.argtypes = [c_void_p, c_int, c_int, ctypes.c_double, c_void_p,
                                            c_void_p, c_void_p]
    lib.bcm_engine_part.argtypes = [c_void_p, c_int, c_int]
    lib.bcm_engine_get_state.argtypes = [c_void_p, c_int, c_void_p]
    lib.bcm_engine_set_state.argtypes = [c_void_p, c_int, c_void_p]
    lib.bcm_engine_get_er.argtypes = [c_void_p, c_void_p]
    lib.bcm_engine_sync.argtypes = [c_void_p]
    lib.bcm_engine_time_part.argtypes = [c_void_p, c_int, c_int, c_int, P(c_double)]
    lib.bcm_engine_ghost_stats.argtypes = [c_void_p, c_void_p]
    lib.bcm_engine_set_ghost_lds_limit.argtypes = [c_void_p, c_int]
    lib.bcm_selftest_division.argtypes = [ctypes.c_ulonglong, c_int, c_void_p]
    lib.bcm_engine_set_flux.argtypes = [c_void_p, c_int, c_int, c_int]
    lib.bcm_engine_autotune.argtypes = [c_void_p, c_void_p]
    lib.bcm_bytes_per_cube.restype = ctypes.c_longlong
    lib.bcm_bytes_per_cube.argtypes = [c_int]
    lib.bcm_engine_alloc_pool.argtypes = [c_void_p, c_int, c_int, c_int]
    lib.bcm_engine_precision.argtypes = [c_void_p]
    lib.bcm_engine_add_case.argtypes = [c_void_p, c_void_p, P(c_int)]
    lib.bcm_engine_remove_case.argtypes = [c_void_p, c_int]
    lib.bcm_engine_case_state.argtypes = [c_void_p, c_int, c_int, c_void_p, c_int]
    lib.bcm_engine_case_er.argtypes = [c_void_p, c_int, c_int, c_void_p]
    lib.bcm_engine_pool_info.argtypes = [c_void_p, c_void_p]
    lib.bcm_case_read.argtypes = [c_char_p, P(CaseParamsC), P(c_void_p)]
    lib.bcm_case_build_on.argtypes = [P(CaseParamsC), c_int, P(c_void_p)]
    lib.bcm_case_read_on.argtypes = [c_char_p, P(CaseParamsC), c_int, P(c_void_p)]
    lib.bcm_case_write_ibm.argtypes = [c_void_p, c_char_p]
    lib.bcm_case_write_plot3d.argtypes = [c_void_p, c_void_p, c_int, c_char_p]
    LL = ctypes.c_longlong
    lib.bcm_write_restart.argtypes = [c_char_p, c_int, LL, c_void_p, c_void_p, c_void_p, c_void_p]
    lib.bcm_read_restart.argtypes = [c_char_p, c_int, P(LL), c_void_p, c_void_p, c_void_p,
                                     c_void_p]
    lib.bcm_engine_output.argtypes = [c_void_p, c_int, c_int, c_char_p]
    lib.bcm_engine_output_wait.argtypes = [c_void_p]
    lib.bcm_engine_save_restart.argtypes = [c_void_p, c_int, c_char_p, LL, c_void_p]
    lib.bcm_engine_load_restart.argtypes = [c_void_p, c_int, c_char_p, P(LL), c_void_p]
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise BCMError(_lib.bcm_last_error().decode())


class Case:
    """A flow case built on the host in the reference's array layouts."""

    def __init__(self, params, device=None, _from_dir=None):
        """Builds the case on the host; with ``device`` (a GPU index) the
        quadratic passes of the IB setup run on that GPU, with bitwise the
        same result."""
        lib = load_library()
        self.params = dict(params)
        p = CaseParamsC()
        for d in range(3):
            p.lattice[d] = int(params.get("lattice", (0, 0, 0))[d])
            p.origin[d] = float(params.get("origin", (0.0, 0.0, 0.0))[d])
            p.sphere_center[d] = float(params.get("sphere_center", (0.0, 0.0, 0.0))[d])
        p.cube_size = float(params.get("cube_size", 0.0))
        p.sphere_diameter = float(params["sphere_diameter"])
        p.ico_level = int(params.get("ico_level", 0))
        p.mach = float(params["mach"])
        p.reynolds = float(params["reynolds"])
        p.seed = int(params.get("seed", 0))
        p.noise_amp = float(params.get("noise_amp", 0.0))
        refine = params.get("refine")
        if refine is not None:
            for d in range(3):
                p.refine_lo[d] = int(refine[0][d])
                p.refine_hi[d] = int(refine[1][d])
        refine2 = params.get("refine2")
        if refine2 is not None:
            for d in range(3):
                p.refine2_lo[d] = int(refine2[0][d])
                p.refine2_hi[d] = int(refine2[1][d])
        h = ctypes.c_void_p()
        if _from_dir is None and device is None:
            _check(lib.bcm_case_build(ctypes.byref(p), ctypes.byref(h)))
        elif _from_dir is None:
            _check(lib.bcm_case_build_on(ctypes.byref(p), int(device), ctypes.byref(h)))
        elif device is None:
            _check(lib.bcm_case_read(os.fsencode(_from_dir), ctypes.byref(p), ctypes.byref(h)))
        else:
            _check(lib.bcm_case_read_on(os.fsencode(_from_dir), ctypes.byref(p), int(device),
                                        ctypes.byref(h)))
        self._h = h
        info = (ctypes.c_int * 16)()
        phys = (ctypes.c_double * 8)()
        _check(lib.bcm_case_info(h, info, phys))
        keys = ("ncube", "n_wallcube", "nbc_minus", "nbc_plus", "nXbc_l", "nXbc_u", "nYbc_l",
                "nYbc_u", "nZbc_l", "nZbc_u", "nadjX_eq", "nadjY_eq", "nadjZ_eq", "n_triangles",
                "n_ghost_candidates")
        self.info = {k: info[n] for n, k in enumerate(keys)}
        self.phys = dict(zip(("rho0", "U0", "Uin", "V0", "W0", "P0", "T0", "mu_L"), list(phys)))
        self.ncube = self.info["ncube"]

    @classmethod
    def from_files(cls, directory, params, device=None):
        """A case from the reference's input files BCMgrid00000.dat and
        BCM_STL_fine.stl in ``directory``.  ``params`` needs ``mach``,
        ``reynolds`` and ``sphere_diameter`` (the reference length of the
        viscosity) and may give ``seed`` and ``noise_amp``.  Rank-decomposition
        files in ``directory`` must describe one rank; a decomposition over
        several ranks raises BCMError."""
        return cls(params, device=device, _from_dir=directory)

    def array(self, name):
        """Numpy view of a case array (owned by the case)."""
        n = ctypes.c_longlong()
        ptr = _lib.bcm_case_array(self._h, name.encode(), ctypes.byref(n))
        if not ptr:
            raise BCMError(_lib.bcm_last_error().decode())
        dbl = name in ("U1_", "U1", "U1q", "cube_size", "Xcube", "Ycube", "Zcube",
                       "Xcnt", "Ycnt", "Zcnt", "ib_minus_w", "ib_plus_w")
        ct = ctypes.c_double if dbl else ctypes.c_int
        buf = (ct * n.value).from_address(ptr)
        a = np.ctypeslib.as_array(buf)
        if name in ("U1_", "U1", "U1q"):
            a = a.reshape(self.ncube, 20, 20, 20, 5)
        elif name in ("FWS", "FWS_grid"):
            a = a.reshape(self.ncube, 20, 20, 20)
        elif name == "adj_number":
            a = a.reshape(self.ncube, 5, 7)
        elif name in ("Xcnt", "Ycnt", "Zcnt"):
            a = a.reshape(self.ncube, 20)
        return a

    def write_inputs(self, directory):
        """BCMgrid00000.dat and BCM_STL_fine.stl, the reference's inputs, and
        its rank-decomposition files MPI_communication_table_in_turn.dat,
        Communication_irank.00000.dat and rank_map.dat for one rank."""
        _check(_lib.bcm_case_write_inputs(self._h, os.fsencode(directory)))

    def write_ibm(self, directory):
        """IBM_minus00000.dat and IBM_plus00000.dat in the reference's format."""
        _check(_lib.bcm_case_write_ibm(self._h, os.fsencode(directory)))

    def write_plot3d(self, directory, step, state=None):
        """PLOT3D BCMdata00001.x and qBCMdata<step>.q of ``state`` (an AoS
        state of this case; default its U1_)."""
        ptr = None
        if state is not None:
            a = np.ascontiguousarray(state, dtype=np.float64)
            if a.size != self.ncube * 40000:
                raise BCMError("state does not match the case's cube count")
            ptr = a.ctypes.data
        _check(_lib.bcm_case_write_plot3d(self._h, ptr, int(step), os.fsencode(directory)))

    def cells(self):
        return (self.ncube - 1) * 4096

    def close(self):
        if getattr(self, "_h", None):
            _lib.bcm_case_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


PRECISIONS = {"fp64": 0, "mixed": 1}

# Area of the reference's force normalisation: er[6] and er[7] are the x- and
# z-forces divided by 0.5 rho0 U0^2 times this constant
# (BCM_Flux_XYZ_Viscous_Runge_kutta.cpp:2637-2638).
FORCE_AREA = 2.602


def selftest_division(n, seed=1):
    """Quotients (of 4 * n on random operands) whose bits differ between the
    FP64 kernels' shared-denominator division and plain division."""
    out = ctypes.c_ulonglong(0)
    _check(load_library().bcm_selftest_division(int(seed), int(n), ctypes.byref(out)))
    return int(out.value)


def drag_coefficient(er6, diameter):
    """Drag coefficient of the sphere of diameter D from the normalised
    x-force er[6]: Cd = -er[6] * 2.602 / (pi D^2 / 4).  The reference prints
    -er[6] as its force coefficient (main.cpp:1284); its area 2.602 is
    replaced by the sphere's frontal area."""
    return -float(er6) * FORCE_AREA / (np.pi * float(diameter) ** 2 / 4.0)


def write_restart(path, step, U1_, U1, U1q, E=None):
    """Restart snapshot of the three time levels (AoS states of one case),
    the step and the norm scales E1..E5."""
    lib = load_library()
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (U1_, U1, U1q)]
    ncube = arrs[0].size // 40000
    if any(a.size != ncube * 40000 for a in arrs):
        raise BCMError("the three states differ in size")
    e = None if E is None else np.ascontiguousarray(E, dtype=np.float64)
    if e is not None and e.size != 5:
        raise BCMError("E holds five norm scales")
    _check(lib.bcm_write_restart(os.fsencode(path), ncube, int(step),
                                 None if e is None else e.ctypes.data,
                                 *(a.ctypes.data for a in arrs)))


def read_restart(path, ncube):
    """(step, E[5], U1_, U1, U1q) of a restart file holding ``ncube`` cubes."""
    lib = load_library()
    step = ctypes.c_longlong()
    e = np.zeros(5)
    arrs = [np.empty((ncube, 20, 20, 20, 5)) for _ in range(3)]
    _check(lib.bcm_read_restart(os.fsencode(path), int(ncube), ctypes.byref(step), e.ctypes.data,
                                *(a.ctypes.data for a in arrs)))
    return (step.value, e, *arrs)


def bytes_per_cube(precision="fp64"):
    """Device memory one pool cube takes."""
    return int(load_library().bcm_bytes_per_cube(PRECISIONS[precision]))


class Engine:
    """Device-resident cases on one GPU: one case (load_case) or a pool of
    independent cases that run together (alloc_pool / add_case)."""

    def __init__(self, device=0):
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib.bcm_engine_create(int(device), ctypes.byref(h)))
        self._h = h
        self.case = None
        self.max_cases = 1
        self.slots = {}

    def load_case(self, case):
        _check(_lib.bcm_engine_load_case(self._h, case._h))
        self.case = case
        self.max_cases = 1
        self.slots = {0: case}

    def alloc_pool(self, cap_cubes, max_cases, precision="fp64"):
        _check(_lib.bcm_engine_alloc_pool(self._h, int(cap_cubes), int(max_cases),
                                          PRECISIONS[precision]))
        self.case = None
        self.max_cases = int(max_cases)
        self.slots = {}

    def add_case(self, case):
        slot = ctypes.c_int()
        _check(_lib.bcm_engine_add_case(self._h, case._h, ctypes.byref(slot)))
        self.slots[slot.value] = case
        return slot.value

    def remove_case(self, slot):
        _check(_lib.bcm_engine_remove_case(self._h, int(slot)))
        self.slots.pop(slot, None)

    def case_state(self, slot, name):
        case = self.slots[slot]
        out = np.empty((case.ncube, 20, 20, 20, 5))
        _check(_lib.bcm_engine_case_state(self._h, int(slot), STATE_NAMES.index(name),
                                          out.ctypes.data, 0))
        return out

    def set_case_state(self, slot, name, arr):
        case = self.slots[slot]
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.size == case.ncube * 40000
        _check(_lib.bcm_engine_case_state(self._h, int(slot), STATE_NAMES.index(name),
                                          a.ctypes.data, 1))

    def case_er(self, slot, inner=0):
        out = np.zeros(8)
        _check(_lib.bcm_engine_case_er(self._h, int(slot), int(inner), out.ctypes.data))
        return out

    def output(self, slot, step, directory):
        """Queues the PLOT3D output of a resident case's U1_ (asynchronous;
        output_wait reports write errors)."""
        _check(_lib.bcm_engine_output(self._h, int(slot), int(step), os.fsencode(directory)))

    def output_wait(self):
        _check(_lib.bcm_engine_output_wait(self._h))

    def save_restart(self, slot, path, step, E=None):
        e = None if E is None else np.ascontiguousarray(E, dtype=np.float64)
        _check(_lib.bcm_engine_save_restart(self._h, int(slot), os.fsencode(path), int(step),
                                            None if e is None else e.ctypes.data))

    def load_restart(self, slot, path):
        """Replaces a resident case's time levels by a restart file's; returns
        (step, E[5])."""
        step = ctypes.c_longlong()
        e = np.zeros(5)
        _check(_lib.bcm_engine_load_restart(self._h, int(slot), os.fsencode(path),
                                            ctypes.byref(step), e.ctypes.data))
        return step.value, e

    def pool_info(self):
        v = (ctypes.c_int * 6)()
        _check(_lib.bcm_engine_pool_info(self._h, v))
        keys = ("cap_cubes", "span_cubes", "active", "max_cases", "free_range", "used_cubes")
        return dict(zip(keys, list(v)))

    def set_time(self, deltaT=1e-4, deltaTau=1e-4 / 200.0, e=0.1):
        _check(_lib.bcm_engine_set_time(self._h, deltaT, deltaTau, e))

    def run(self, n_steps, n_inner, record_er=True, use_graph=True):
        """Runs every resident case.  Returns er[1..7] per inner iteration:
        shape (n_steps*n_inner, 7) for a single case, (n_steps*n_inner,
        max_cases, 7) for a pool."""
        mc = self.max_cases
        er = np.zeros((max(n_steps, 1) * n_inner, mc, 7)) if record_er else None
        ptr = er.ctypes.data if record_er else None
        _check(_lib.bcm_engine_run(self._h, int(n_steps), int(n_inner), ptr, int(bool(use_graph))))
        if not record_er:
            return None
        er = er[: n_steps * n_inner]
        return er[:, 0, :] if self.case is not None else er

    def run_converge(self, n_steps, max_inner, tol=0.02, E=None, record_er=True):
        """Steps under the reference's inner-iteration control
        (main.cpp:1249-1320): each case's inner loop ends when its five
        running-max normalised norms are below tol, or after max_inner
        iterations.  E: running maxima E1..E5 per case slot (zeros, the
        reference's start, when None).  Returns (iterations per step, norms
        er[1..7] per inner iteration with zero rows after each break, E);
        per-slot axes are dropped for a single case."""
        mc = self.max_cases
        E = np.zeros((mc, 5)) if E is None else np.array(E, dtype=np.float64).reshape(mc, 5).copy()
        iters = np.zeros((max(n_steps, 1), mc), dtype=np.int32)
        er = np.zeros((max(n_steps, 1) * max_inner, mc, 7)) if record_er else None
        _check(_lib.bcm_engine_run_converge(self._h, int(n_steps), int(max_inner), float(tol),
                                            E.ctypes.data, iters.ctypes.data,
                                            er.ctypes.data if record_er else None))
        iters = iters[:n_steps]
        if record_er:
            er = er[: n_steps * max_inner]
        if self.case is not None:
            return iters[:, 0], (er[:, 0, :] if record_er else None), E[0]
        return iters, er, E

    def part(self, name, RK=1):
        _check(_lib.bcm_engine_part(self._h, PARTS[name], int(RK)))

    def get_state(self, name):
        out = np.empty((self.case.ncube, 20, 20, 20, 5))
        _check(_lib.bcm_engine_get_state(self._h, STATE_NAMES.index(name), out.ctypes.data))
        return out

    def set_state(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.size == self.case.ncube * 40000
        _check(_lib.bcm_engine_set_state(self._h, STATE_NAMES.index(name), a.ctypes.data))

    def er(self):
        out = np.zeros(8)
        _check(_lib.bcm_engine_get_er(self._h, out.ctypes.data))
        return out

    def sync(self):
        _check(_lib.bcm_engine_sync(self._h))

    def time_part(self, name, RK=1, reps=10):
        ms = ctypes.c_double()
        _check(_lib.bcm_engine_time_part(self._h, PARTS[name], int(RK), int(reps), ctypes.byref(ms)))
        return ms.value

    def set_flux(self, variant, nseg=1, fuse=1):
        """Pins the flux kernel choice (no autotuning).  FP64: fuse 0
        separate update, 1 update fused in place, 2 update and face halos
        written to a second copy of the iterate (engine stages).  Mixed: only
        nseg counts, the runs per cube (1, 2 or 4) of its one kernel."""
        _check(_lib.bcm_engine_set_flux(self._h, int(variant), int(nseg), int(fuse)))

    def autotune(self):
        out = (ctypes.c_int * 3)()
        _check(_lib.bcm_engine_autotune(self._h, out))
        return {"variant": out[0], "nseg": out[1], "fuse": int(out[2])}

    def set_ghost_lds_limit(self, slots):
        """Cubes whose IB fill needs more LDS value slots than this take the
        in-place schedule on memory (cases loaded or added afterwards); the
        default 1536 is the kernel's LDS capacity.  A test hook."""
        _check(_lib.bcm_engine_set_ghost_lds_limit(self._h, int(slots)))

    def ghost_stats(self):
        s = (ctypes.c_int * 3)()
        _check(_lib.bcm_engine_ghost_stats(self._h, s))
        return {"entries": s[0], "cubes": s[1], "levels": s[2]}

    def close(self):
        if getattr(self, "_h", None):
            _lib.bcm_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
