"""Flow cases of the benchmark and test suite.

All cases: a sphere in a box of 16^3-cell cubes (one cube size, or two sizes
for the two-level grids),
gamma = 1.4, Pr = 0.72, freestream density 1.1842 and pressure 101300 (the
reference's prm.h values), seeded +-1e-3 relative density noise on fluid cells.
"""

_BASE = dict(origin=(0.0, 0.0, 0.0), sphere_diameter=0.25, seed=20260501, noise_amp=1e-3)

CONFIGS = {
    # 2x2x2 cubes, sphere across all eight cubes: golden-fixture and unit-test size.
    "tiny": dict(_BASE, lattice=(2, 2, 2), cube_size=0.25, sphere_center=(0.25, 0.25, 0.25),
                 ico_level=4, mach=0.2, reynolds=100.0),
    # 4x4x4 cubes on the unit box (262,144 cells), level-4 icosphere, M 0.2, Re 100.
    "cfg1": dict(_BASE, lattice=(4, 4, 4), cube_size=0.25, sphere_center=(0.5, 0.5, 0.5),
                 ico_level=4, mach=0.2, reynolds=100.0),
    # 16x16x8 cubes of edge 1/16 (box 1 x 1 x 0.5, 8,388,608 cells), level-5
    # icosphere at the box centre, M 0.3, Re 300.
    "cfg2": dict(_BASE, lattice=(16, 16, 8), cube_size=1.0 / 16, sphere_center=(0.5, 0.5, 0.25),
                 ico_level=5, mach=0.3, reynolds=300.0),
    # Two-level grids (cubes of two sizes): a 2x2x2 lattice of cubes of edge 0.5
    # with the first cube split into 8 (15 cubes), the sphere inside the split
    # cube; and a 4x4x4 lattice of edge 0.25 with the central 2x2x2 split
    # (120 cubes), the sphere inside the refined block.
    "ml_tiny": dict(_BASE, lattice=(2, 2, 2), cube_size=0.5, sphere_center=(0.25, 0.25, 0.25),
                    sphere_diameter=0.2, ico_level=4, mach=0.2, reynolds=100.0,
                    refine=((0, 0, 0), (1, 1, 1))),
    "ml": dict(_BASE, lattice=(4, 4, 4), cube_size=0.25, sphere_center=(0.5, 0.5, 0.5),
               ico_level=4, mach=0.2, reynolds=100.0, refine=((1, 1, 1), (3, 3, 3))),
    # Three-level grid: the ml lattice with the central 2x2x2 of its half-size
    # cubes (doubled-lattice indices 3..4) split again (176 cubes of edges 0.25,
    # 0.125 and 0.0625), a sphere of diameter 0.15 inside the finest block.
    "ml3": dict(_BASE, lattice=(4, 4, 4), cube_size=0.25, sphere_center=(0.5, 0.5, 0.5),
                sphere_diameter=0.15, ico_level=4, mach=0.2, reynolds=100.0,
                refine=((1, 1, 1), (3, 3, 3)), refine2=((3, 3, 3), (5, 5, 5))),
}

# Stepping of each case: (physical steps, inner iterations per step).
STEPPING = {"tiny": (2, 3), "cfg1": (20, 10), "cfg2": (200, 10), "batch": (50, 10),
            "ml_tiny": (2, 3), "ml": (4, 5), "ml3": (3, 4)}

# Runs under the reference's inner-iteration control (main.cpp:1249-1320):
# physical steps, the iteration cap (iteration_end_step) and the tolerance on
# the running-max normalised norms.  The reference's tolerance is 0.02; 0.97
# makes these short runs end their inner loops at different iterations.
CONVERGE = {"tiny": (8, 12, 0.97), "ml_tiny": (8, 12, 0.97), "cfg1": (12, 10, 0.97), "ml": (12, 10, 0.97)}

# Cube lattices of the batched workload: n^3 cubes on the unit box.
BATCH_LATTICES = (2, 3, 4, 5, 6, 8)


def _rng(seed):
    import numpy as np
    return np.random.default_rng(seed)


def batch_cases(n_cases=64, seed=20260502, lattices=BATCH_LATTICES, ico_level=4):
    """Independent sphere cases of the batched workload, drawn from seed:
    an n^3 lattice of cubes on the unit box with n uniform in lattices,
    M uniform in [0.1, 0.5], Re uniform in [50, 1000] and the sphere centre
    offset from the box centre by up to +-0.1 per axis.  Each case gets its
    own density-noise seed."""
    r = _rng(seed)
    out = []
    for c in range(n_cases):
        n = int(r.choice(lattices))
        centre = tuple(float(0.5 + r.uniform(-0.1, 0.1)) for _ in range(3))
        out.append(dict(_BASE, lattice=(n, n, n), cube_size=1.0 / n, sphere_center=centre,
                        ico_level=ico_level, mach=float(r.uniform(0.1, 0.5)),
                        reynolds=float(r.uniform(50.0, 1000.0)), seed=int(seed * 1000 + c)))
    return out


def case_params(name, **overrides):
    p = dict(CONFIGS[name])
    p.update(overrides)
    return p
