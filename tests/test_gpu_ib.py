"""GPU tests of the device IB setup: the triangle filter and the surface
search of BCM_Immersed_boundary as kernels.

Bar: the case built with the search on the GPU is bitwise the host-built
case in every array (flags, IB tables, initial state, lists), which the
setup fixtures pin to the reference.  Cases: the fixture cases, cfg2 (level-5
surface), the edge cases, a batch of random cases, and surfaces altered so
that rays miss (dropped facets), hit back faces (reversed facets) or meet
degenerate facets, which exercises the closest-point fallback and the
not-a-number paths of the ray test.
"""
import time

import numpy as np
import pytest

from bcm_cgli_amd.configs import batch_cases

pytestmark = pytest.mark.gpu

ARRAYS = ("FWS", "FWS_grid", "U1_", "U1", "U1q", "csl", "ib_minus_gc", "ib_minus_ip",
          "ib_minus_w", "ib_plus_gc", "ib_plus_ip", "ib_plus_w", "Xbc_l", "Xbc_u", "Ybc_l",
          "Ybc_u", "Zbc_l", "Zbc_u", "adjX_eq", "adjY_eq", "adjZ_eq")

EDGE = {
    "single_cube": dict(lattice=(1, 1, 1), cube_size=0.5, sphere_center=(0.25, 0.25, 0.25)),
    "no_body": dict(lattice=(2, 1, 1), sphere_diameter=0.0, sphere_center=(0.1, 0.1, 0.1)),
    "ragged_cut": dict(lattice=(3, 2, 1), sphere_center=(0.4, 0.25, 0.05)),
    "inside_one_cube": dict(sphere_center=(0.125, 0.375, 0.125), sphere_diameter=0.15,
                            ico_level=2),
}


def assert_same_case(a, b, label):
    assert a.info == b.info, label
    for n in ARRAYS:
        x, y = a.array(n), b.array(n)
        assert x.shape == y.shape, (label, n)
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (label, n)


@pytest.mark.parametrize("name", ["tiny", "cfg1", "cfg2"])
def test_device_setup_matches_host(B, name):
    p = B.case_params(name)
    t0 = time.perf_counter()
    host = B.Case(p)
    t1 = time.perf_counter()
    dev = B.Case(p, device=0)
    t2 = time.perf_counter()
    print(f"{name}: host setup {t1 - t0:.3f} s, device setup {t2 - t1:.3f} s, "
          f"{dev.info['n_ghost_candidates']} ghost cells x {dev.info['n_triangles']} triangles")
    assert_same_case(host, dev, name)


@pytest.mark.parametrize("label", sorted(EDGE))
def test_device_setup_edge_cases(B, label):
    p = B.case_params("tiny", **EDGE[label])
    assert_same_case(B.Case(p), B.Case(p, device=0), label)


def test_device_setup_batch(B):
    for n, p in enumerate(batch_cases(6, seed=99, lattices=(2, 3, 4), ico_level=3)):
        assert_same_case(B.Case(p), B.Case(p, device=0), f"batch {n}")


def _alter_stl(path, degenerate):
    lines = open(path).read().split("\n")
    head, body = lines[0], lines[1:]
    facets, rest, n = [], [], 0
    while n < len(body):
        if body[n].startswith("facet normal"):
            facets.append(body[n:n + 7])
            n += 7
        else:
            rest.append(body[n])
            n += 1
    out = []
    # Optionally a zero-area facet first (its not-a-number ray test counts as
    # a hit for every ghost cell), then every third facet dropped and every
    # fifth facet's vertices reversed (a back face for the ray test).
    if degenerate:
        v = facets[0][2]
        out.append(["facet normal 0 0 0", "outer loop", v, v, v, "endloop", "endfacet"])
    for i, f in enumerate(facets):
        if i % 3 == 2:
            continue
        if i % 5 == 0:
            f = f[:2] + [f[4], f[3], f[2]] + f[5:]
        out.append(f)
    open(path, "w").write("\n".join([head] + [x for f in out for x in f] + rest))
    return len(out)


@pytest.mark.parametrize("degenerate", [False, True])
def test_device_setup_with_altered_surface(B, tmp_path, degenerate):
    p = B.case_params("cfg1")
    B.Case(p).write_inputs(str(tmp_path))
    nfacets = _alter_stl(str(tmp_path / "BCM_STL_fine.stl"), degenerate)
    host = B.Case.from_files(str(tmp_path), p)
    dev = B.Case.from_files(str(tmp_path), p, device=0)
    assert host.info["n_triangles"] == nfacets
    print(f"altered surface: {nfacets} facets, IB {host.info['nbc_minus']} minus / "
          f"{host.info['nbc_plus']} plus of {host.info['n_ghost_candidates']} ghost cells")
    assert_same_case(host, dev, "altered")
