"""The drop-in boundary as a compiled C++ caller sees it: a translation unit
shaped like the reference's main program (tests/cpp/main_shaped.cpp: the
Array.h array types and the argument lists of main.cpp:562-572,
main.cpp:1157-1245 and main.cpp:1346-1351) compiles against
include/bcm_cgli.hpp with warnings as errors and links against
libbcm_cgli.so; the C header alone is valid C.  Nothing is run (no device)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

INC = os.path.join(ROOT, "include")
SRC = os.path.join(ROOT, "tests", "cpp", "main_shaped.cpp")


def _cxx():
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no g++")
    return cxx


def test_main_shaped_caller_compiles_and_links(B, tmp_path):
    libdir = os.path.dirname(B.LIB_PATH)
    out = tmp_path / "main_shaped"
    r = subprocess.run([_cxx(), "-std=c++11", "-Wall", "-Wextra", "-Werror", f"-I{INC}", SRC,
                        f"-L{libdir}", "-lbcm_cgli", "-Wl,-rpath-link,/opt/rocm/lib",
                        f"-Wl,-rpath,{libdir}", "-o", str(out)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert out.exists()


def test_flat_pointer_calls_need_the_overloads(tmp_path):
    # Without bcm_cgli.hpp the Array.h types do not convert to the C ABI's
    # flat pointers: the overloads are what makes the caller compile.
    src = tmp_path / "flat.cpp"
    src.write_text('#include "bcm_cgli.h"\n'
                   "int main() {\n"
                   "  int (*adj_number)[5][7] = new int[9][5][7];\n"
                   "  int n[15];\n  int* l[15];\n  int* csl = new int[9];\n"
                   "  BCM_Interface_table(0, 9, 1, &n[0], &n[1], &n[2], &n[3], &n[4], &n[5], &n[6],\n"
                   "    &n[7], &n[8], &n[9], &n[10], &n[11], &n[12], &n[13], &n[14], csl,\n"
                   "    adj_number, l[0], l[1], l[2], l[3], l[4], l[5], l[6], l[7], l[8], l[9],\n"
                   "    l[10], l[11], l[12], l[13], l[14]);\n"
                   "  return 0;\n}\n")
    r = subprocess.run([_cxx(), "-std=c++11", "-fsyntax-only", f"-I{INC}", str(src)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    hpp = tmp_path / "hpp.cpp"
    hpp.write_text(src.read_text().replace('"bcm_cgli.h"', '"bcm_cgli.hpp"'))
    r = subprocess.run([_cxx(), "-std=c++11", "-fsyntax-only", f"-I{INC}", str(hpp)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]


def test_c_header_is_c(tmp_path):
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("no gcc")
    src = tmp_path / "c.c"
    src.write_text('#include "bcm_cgli.h"\nint main(void) { return bcm_last_error() == 0; }\n')
    r = subprocess.run([cc, "-std=c99", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", f"-I{INC}",
                        str(src)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
