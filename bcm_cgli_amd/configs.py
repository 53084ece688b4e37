"""Flow cases of the benchmark and test suite.

All cases: a sphere of diameter D = 0.25 in a box of uniform 16^3-cell cubes,
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
}

# Stepping of each case: (physical steps, inner iterations per step).
STEPPING = {"tiny": (2, 3), "cfg1": (20, 10), "cfg2": (200, 10)}


def case_params(name, **overrides):
    p = dict(CONFIGS[name])
    p.update(overrides)
    return p
