"""Shared helpers of the test suite."""
import hashlib
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(arr):
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()


def golden(name):
    with open(os.path.join(GOLDEN, f"{name}_golden.json")) as f:
        g = json.load(f)
    g["er"] = np.array([[float.fromhex(x) for x in row] for row in g["er_hex"]])
    return g


def golden_converge(name):
    """Fixture of a run under the reference's convergence break: iterations per
    step, final-array hashes and the raw norms er[1..7] of every iteration run,
    laid out as (steps * max_inner, 7) with zero rows after each break."""
    with open(os.path.join(GOLDEN, f"{name}_converge_golden.json")) as f:
        g = json.load(f)
    er = np.zeros((g["n_steps"] * g["max_inner"], 7))
    for row in g["er_hex"]:
        er[(row[0] - 1) * g["max_inner"] + row[1] - 1] = [float.fromhex(x) for x in row[2:]]
    g["er"] = er
    return g


def golden_ib(name):
    return np.load(os.path.join(GOLDEN, f"{name}_ib.npz"))


def bitwise_mismatch(a, b):
    """Number of elements whose bit patterns differ."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return int((a.view(np.int64) != b.view(np.int64)).sum())


def describe_mismatch(a, b):
    d = a.view(np.int64) != b.view(np.int64)
    idx = np.argwhere(d)
    return f"{d.sum()} differing values, first at {idx[:3].tolist()}, max abs diff {np.abs(a - b).max():.3e}"
