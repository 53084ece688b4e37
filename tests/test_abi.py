"""The C ABI library loads without a GPU and exports every declared entry point."""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "bcm_cgli.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    names = set(re.findall(r"\b((?:bcm|BCM)_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_reference_dropins():
    names = declared_functions()
    for ref in ("BCM_X_boundary_condition", "BCM_Y_boundary_condition", "BCM_Z_boundary_condition",
                "BCM_Ghostcell_minus", "BCM_Ghostcell_plus", "BCM_Interface_table", "BCM_Interface",
                "BCM_Flux_XYZ_Viscous_Runge_kutta", "BCM_Grid", "BCM_Reading_IBM", "BCM_Output"):
        assert ref in names


def test_library_exports_every_declared_symbol(B):
    lib = ctypes.CDLL(B.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_engine_create_without_device_fails_cleanly(B):
    # In a container without a GPU, engine creation must report an error
    # rather than fall back to a CPU path.
    import bcm_cgli_amd
    lib = bcm_cgli_amd.load_library()
    h = ctypes.c_void_p()
    rc = lib.bcm_engine_create(99, ctypes.byref(h))
    assert rc != 0
    assert lib.bcm_last_error()
