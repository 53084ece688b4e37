"""The flux/RK stage entry point that allocates nothing (bcm_flux_rk_noalloc)
against the drop-in BCM_Flux_XYZ_Viscous_Runge_kutta, through the C ABI.

The no-alloc entry point runs on device arrays the caller owns (allocated
here with the HIP runtime, in the library's device layout) and keeps its
norm partials and case tables in a fixed device array that every call
reuses.  Bar: bitwise equal arrays and norms for a stage sequence RK 1, 2, 3
on a 64-cube case, then on an 8-cube case (a second, smaller input in the
same process, so the second call sees the first call's scratch contents).
"""
import ctypes

import numpy as np
import pytest

from helpers import describe_mismatch

pytestmark = pytest.mark.gpu

STATE = ("U1_", "U1", "U1q", "U1p1", "U1p2", "Rku1", "Residual1")
OUT = ("U1_", "U1p1", "U1p2", "Rku1", "Residual1")
DT, DTAU, EPS = 1e-4, 1e-4 / 200.0, 0.1


def hip_runtime():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            lib = ctypes.CDLL(name)
            break
        except OSError:
            continue
    else:
        pytest.fail("HIP runtime library not found")
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    lib.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
    lib.hipFree.argtypes = [vp]
    lib.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
    lib.hipSetDevice.argtypes = [ctypes.c_int]
    return lib


class DeviceArrays:
    """Device copies of host arrays, allocated by the caller (this test)."""

    def __init__(self, hip):
        self.hip = hip
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def get(self, p, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def soa(a):
    """[ncube][20][20][20][5] -> [ncube][5][20][20][20] (the device layout)."""
    return np.ascontiguousarray(np.moveaxis(a, 4, 1))


def aos(s):
    return np.ascontiguousarray(np.moveaxis(s, 1, 4))


def host_state(case):
    nc = case.ncube
    st = {n: np.zeros((nc, 20, 20, 20, 5)) for n in STATE}
    for n in ("U1_", "U1", "U1q"):
        st[n][...] = case.array(n)
    return st


def phys7(case):
    ph = case.phys
    return np.array([ph["rho0"], ph["U0"], ph["Uin"], ph["V0"], ph["W0"], ph["P0"], ph["mu_L"]])


def dropin_stages(B, case):
    """RK 1, 2, 3 through the drop-in on host arrays bound to an engine."""
    lib = B.load_library()
    P, c_int, c_double = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    nc = case.ncube
    st = host_state(case)
    fws = np.ascontiguousarray(case.array("FWS"), dtype=np.int32)
    cs = np.ascontiguousarray(case.array("cube_size"))
    csl = np.ascontiguousarray(case.array("csl"))
    p7 = phys7(case)
    h = ctypes.c_void_p()
    assert lib.bcm_engine_create(0, ctypes.byref(h)) == 0
    try:
        lib.bcm_bind.argtypes = [P, c_int, c_int] + [P] * 10
        assert lib.bcm_bind(h, nc, nc - 1, *[st[n].ctypes.data for n in STATE], fws.ctypes.data,
                            cs.ctypes.data, p7.ctypes.data) == 0, lib.bcm_last_error()
        lib.BCM_Flux_XYZ_Viscous_Runge_kutta.argtypes = [c_int, c_int, c_int, c_double, c_double,
                                                         c_double] + [P] * 11
        er = np.zeros(8)
        for RK in (1, 2, 3):
            lib.BCM_Flux_XYZ_Viscous_Runge_kutta(0, nc, RK, DT, DTAU, EPS, fws.ctypes.data,
                                                 csl.ctypes.data, cs.ctypes.data,
                                                 *[st[n].ctypes.data for n in STATE], er.ctypes.data)
        assert lib.bcm_sync_to_host(h) == 0
    finally:
        lib.bcm_engine_destroy(h)
    return st, er


def noalloc_stages(B, hip, case):
    """RK 1, 2, 3 through bcm_flux_rk_noalloc on caller-owned device arrays."""
    lib = B.load_library()
    P, c_int, c_double = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.bcm_flux_rk_noalloc.argtypes = [c_int, c_int, c_int, c_double, c_double, c_double] + [P] * 12
    nc = case.ncube
    st = host_state(case)
    dev = DeviceArrays(hip)
    try:
        d = {n: dev.put(soa(st[n])) for n in STATE}
        fws = dev.put(np.ascontiguousarray(case.array("FWS"), dtype=np.int8))
        cs = dev.put(np.ascontiguousarray(case.array("cube_size")))
        er_d = dev.put(np.full(8, np.nan))
        p7 = phys7(case)
        for RK in (1, 2, 3):
            rc = lib.bcm_flux_rk_noalloc(nc, nc - 1, RK, DT, DTAU, EPS, fws, cs,
                                         *[d[n] for n in STATE], er_d, p7.ctypes.data, None)
            assert rc == 0, lib.bcm_last_error()
        assert hip.hipDeviceSynchronize() == 0
        out = {n: aos(dev.get(d[n], (nc, 5, 20, 20, 20))) for n in OUT}
        er = dev.get(er_d, (8,))
    finally:
        dev.free()
    return out, er


def bitwise(a, b, what):
    if not np.array_equal(a.view(np.int64), b.view(np.int64)):
        raise AssertionError(f"{what}: {describe_mismatch(a, b)}")


def test_noalloc_matches_dropin_two_calls(B):
    hip = hip_runtime()
    assert hip.hipSetDevice(0) == 0
    for name in ("cfg1", "tiny"):    # 64 cubes, then 8 cubes
        case = B.Case(B.case_params(name))
        print(f"no-alloc vs drop-in: {name}, {case.ncube - 1} cubes", flush=True)
        ref, er_ref = dropin_stages(B, case)
        out, er = noalloc_stages(B, hip, case)
        for n in OUT:
            bitwise(out[n], ref[n], f"{name} {n}")
        bitwise(er[1:8], er_ref[1:8], f"{name} er")
        assert er[0] == 0.0


def test_noalloc_rejects_oversize_and_bad_stage(B):
    lib = B.load_library()
    P, c_int, c_double = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.bcm_flux_rk_noalloc.argtypes = [c_int, c_int, c_int, c_double, c_double, c_double] + [P] * 12
    dummy = np.zeros(8)
    ptrs = [dummy.ctypes.data] * 11 + [None]
    big = 16384 + 2
    assert lib.bcm_flux_rk_noalloc(big, big - 1, 1, DT, DTAU, EPS, *ptrs) != 0
    assert b"16384" in lib.bcm_last_error()
    assert lib.bcm_flux_rk_noalloc(9, 8, 4, DT, DTAU, EPS, *ptrs) != 0
    assert b"RK" in lib.bcm_last_error()
