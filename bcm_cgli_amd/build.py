"""Builds the in-tree native libraries.

* ``bcm_cgli_amd/libbcm_cgli.so`` -- the product: HIP kernels for gfx950, the
  device engine, host case setup and the C ABI of ``include/bcm_cgli.h``.
* ``oracle/liboracle.so`` -- the FP64 CPU oracle (test infrastructure).

Both are compiled every time :func:`build` runs.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libbcm_cgli.so")
ORACLE_SRC = os.path.join(ROOT, "oracle", "bcm_oracle.c")
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")

ARCH = os.environ.get("BCM_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: no multiply-add fusion, so every expression rounds exactly
# as the reference's FP64 CPU path does.  Packed FP32 instructions are
# disabled for all device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

SOURCES = ["bcm_kernels.hip", "bcm_kernels_mixed.hip", "bcm_ib.hip", "bcm_noalloc.hip",
           "bcm_engine.cpp", "bcm_output.cpp", "bcm_capi.cpp", "bcm_setup.cpp", "bcm_io.cpp"]
# Sources that include the HIP runtime; the others are plain host C++.
HIP_HOST = ("bcm_engine.cpp", "bcm_output.cpp", "bcm_capi.cpp", "bcm_setup.cpp")

# The mixed-precision kernels carry no bitwise contract with the reference:
# they let the compiler fuse multiply-adds.
PER_FILE_FLAGS = {"bcm_kernels_mixed.hip": ["-ffp-contract=fast"]}


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build step failed: {cmd[0]} (exit {r.returncode})")
    return r.stdout


def build_product(verbose=False, extra=()):
    objs = []
    hipcc = _hipcc()
    build_dir = os.path.join(PKG, "_build")
    os.makedirs(build_dir, exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") or src in HIP_HOST else []
        cmd = [hipcc, *HIP_FLAGS, *PER_FILE_FLAGS.get(src, []), *extra, *lang, "-c",
               os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed = True
    if failed:
        raise RuntimeError("compilation failed")
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-pthread"], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose=False):
    cc = shutil.which("gcc") or "cc"
    tmp = ORACLE_LIB + ".tmp"
    _run([cc, "-O2", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared", "-fPIC",
          "-o", tmp, ORACLE_SRC, "-lm"], verbose)
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build(verbose=False):
    build_oracle(verbose)
    return build_product(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB, ORACLE_LIB)
