"""Writes the per-case configuration headers for the reference build.

Test infrastructure only.  The reference takes its problem size and flow
constants from two compile-time headers, Resolution.h (cube count ``MPI_Ncube``
and rank count ``np``, Resolution.h:5-7) and prm.h (``mu_L``, ``U0``, ``Uin``,
prm.h:8,26-28).  This script reads both from the reference tree, substitutes
those five values for one flow case and writes the result into a build
directory under oracle/_ref/, which the compiler searches before the
reference directory (``-I<cfg> -I-``).  Nothing else in either header changes.

Usage: gen_ref_cfg.py <reference_dir> <out_dir> <n_cubes> <U0> <mu_L>
"""
import os
import re
import sys


def substitute(text, pattern, repl, count=1):
    new, n = re.subn(pattern, repl, text, count=count, flags=re.M)
    if n != count:
        raise SystemExit(f"pattern {pattern!r} matched {n} times, expected {count}")
    return new


def main():
    ref, out, ncubes, u0, mul = sys.argv[1:6]
    os.makedirs(out, exist_ok=True)
    with open(os.path.join(ref, "Resolution.h")) as f:
        res = f.read()
    res = substitute(res, r"^#define np\s+\d+", "#define np 1")
    res = substitute(res, r"^#define MPI_Ncube\s+\d+", f"#define MPI_Ncube {int(ncubes)}")
    with open(os.path.join(out, "Resolution.h"), "w") as f:
        f.write(res)
    with open(os.path.join(ref, "prm.h")) as f:
        prm = f.read()
    prm = substitute(prm, r"^#define mu_L\s+\S+", f"#define mu_L {mul}")
    prm = substitute(prm, r"^double U0 = [^;]+;", f"double U0 = {u0};")
    prm = substitute(prm, r"^double Uin = [^;]+;", f"double Uin = {u0};")
    with open(os.path.join(out, "prm.h"), "w") as f:
        f.write(prm)


if __name__ == "__main__":
    main()
