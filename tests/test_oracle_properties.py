"""Closed-form and symmetry properties of the oracle's individual routines.

These pin each routine independently of the golden fixtures: a dropped term,
a wrong sign or index, or a transposed operand in the flux, viscous,
preconditioner, ghost-fill or halo code breaks one of them.
"""
import numpy as np
import pytest


def single_cube_case(B, **kw):
    p = B.case_params("tiny", lattice=(1, 1, 1), sphere_diameter=0.0, noise_amp=0.0,
                      sphere_center=(0.1, 0.1, 0.1), **kw)
    return B.Case(p)


def smooth_field(case, seed=1, amp=1e-2):
    """Freestream plus a smooth random perturbation on every cell of cube 1."""
    rng = np.random.default_rng(seed)
    ph = case.phys
    x = np.arange(20) / 20.0
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    modes = [rng.normal(size=4) for _ in range(6)]

    def f(m):
        a, b, c, d = modes[m]
        return np.sin(2.1 * X * a + 1.3 * Y * b + 0.7 * Z * c + d)

    rho = ph["rho0"] * (1 + amp * f(0))
    u = ph["U0"] * (1.0 + amp * f(1))
    v = ph["U0"] * (0.3 + amp * f(2))
    w = ph["U0"] * (-0.2 + amp * f(3))
    P = ph["P0"] * (1 + amp * f(4))
    U = np.zeros((case.ncube, 20, 20, 20, 5))
    U[1, ..., 0] = rho
    U[1, ..., 1] = rho * u
    U[1, ..., 2] = rho * v
    U[1, ..., 3] = rho * w
    U[1, ..., 4] = P / 0.4 + 0.5 * rho * (u * u + v * v + w * w)
    return U


def flux_residual(O, case, U):
    orc = O.OracleRun(case, nthreads=1)
    for n in ("U1_", "U1", "U1q"):
        orc.state[n][...] = U
    orc.flux_update(1)
    return orc.state["Rku1"].copy()


def permute_cyclic(U):
    # new (X, Y, Z) takes old (x, y, z) = (Y, Z, X): old z-axis becomes new x
    V = np.transpose(U, (0, 3, 1, 2, 4)).copy()
    out = V.copy()
    out[..., 1] = V[..., 3]
    out[..., 2] = V[..., 1]
    out[..., 3] = V[..., 2]
    return out


def mirror_x(U):
    V = U[:, ::-1].copy()
    V[..., 1] = -V[..., 1]
    return V


def test_freestream_is_preserved_to_roundoff(B, O):
    # Uniform flow, no body: every face sees equal left and right states, so
    # the flux differences, viscous terms and dual-time source vanish up to
    # the round-off of the inflow boundary's conserved/primitive round trip.
    case = B.Case(B.case_params("tiny", lattice=(2, 1, 1), sphere_diameter=0.0, noise_amp=0.0,
                                sphere_center=(0.1, 0.1, 0.1)))
    orc = O.OracleRun(case, nthreads=2)
    U0 = orc.state["U1_"].copy()
    hist = orc.run(1, 2)
    scale = np.array([1.2, 82.0, 82.0, 82.0, 2.6e5])
    d = np.abs(orc.state["U1_"][1:, 2:18, 2:18, 2:18] - U0[1:, 2:18, 2:18, 2:18])
    assert np.all(d.max(axis=(0, 1, 2, 3)) <= 1e-13 * scale)
    assert np.all(np.abs(hist[:, 5:]) == 0)  # no solid cells, no force


def test_residual_is_rotation_equivariant(B, O):
    case = single_cube_case(B)
    U = smooth_field(case)
    R = flux_residual(O, case, U)
    Rp = flux_residual(O, case, permute_cyclic(U))
    interior = (slice(1, 2), slice(2, 18), slice(2, 18), slice(2, 18))
    a = permute_cyclic(R)[interior]
    b = Rp[interior]
    scale = np.abs(a).max(axis=(0, 1, 2, 3))
    assert np.all(scale > 0)
    assert np.all(np.abs(a - b).max(axis=(0, 1, 2, 3)) <= 1e-9 * scale)


def test_residual_is_mirror_equivariant(B, O):
    case = single_cube_case(B)
    U = smooth_field(case, seed=3)
    R = flux_residual(O, case, U)
    Rm = flux_residual(O, case, mirror_x(U))
    interior = (slice(1, 2), slice(2, 18), slice(2, 18), slice(2, 18))
    a = mirror_x(R)[interior]
    b = Rm[interior]
    scale = np.abs(a).max(axis=(0, 1, 2, 3))
    assert np.all(np.abs(a - b).max(axis=(0, 1, 2, 3)) <= 1e-9 * scale)


def test_ghost_fill_of_uniform_field(B, O):
    case = B.Case(B.case_params("tiny"))
    orc = O.OracleRun(case, nthreads=1)
    ph = case.phys
    U = orc.state["U1_"]
    U[1:] = 0
    U[1:, ..., 0] = ph["rho0"]
    U[1:, ..., 1] = ph["rho0"] * ph["U0"]
    U[1:, ..., 4] = ph["P0"] / 0.4 + 0.5 * ph["rho0"] * ph["U0"] ** 2
    orc.ghost()
    gc = case.array("ib_minus_gc")[1:].reshape(-1, 4)
    ip = case.array("ib_minus_ip")[1:].reshape(-1, 4)
    w = case.array("ib_minus_w")[1:].reshape(-1, 8)
    g = U[gc[:, 0], gc[:, 1], gc[:, 2], gc[:, 3]]
    assert np.allclose(g[:, 0], ph["rho0"], rtol=1e-14, atol=0)
    corner = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [1, 0, 1], [0, 1, 1], [1, 1, 1]])
    written = set()
    checked = 0
    for n in range(len(gc)):
        # The fill is in place and in list order: only entries whose corners
        # no earlier entry has overwritten see the uniform field.
        corners = {(ip[n, 0], *(ip[n, 1:] + corner[m])) for m in range(8)}
        fresh = not (corners & written)
        written.add(tuple(gc[n]))
        if not fresh:
            continue
        checked += 1
        rel = gc[n, 1:] - ip[n, 1:]
        self_corner = [m for m in range(8) if np.array_equal(corner[m], rel)]
        if self_corner:
            s = self_corner[0]
            expect_u = ph["U0"] * (w[n].sum() - w[n, s]) / (2 - w[n, s])
        else:
            expect_u = 0.5 * ph["U0"] * w[n].sum()
        assert g[n, 1] == pytest.approx(g[n, 0] * expect_u, rel=1e-12)
        assert g[n, 2] == 0 and g[n, 3] == 0
    assert checked > len(gc) // 4


def test_interface_copies_and_edge_means(B, O):
    case = B.Case(B.case_params("tiny", lattice=(2, 2, 1), sphere_diameter=0.0,
                                sphere_center=(0.1, 0.1, 0.1)))
    orc = O.OracleRun(case, nthreads=1)
    rng = np.random.default_rng(7)
    U = orc.state["U1_"]
    U[1:] = rng.normal(size=U[1:].shape)
    before = U.copy()
    orc.interface()
    adj = case.array("adj_number")
    for c in range(1, case.ncube):
        xp = adj[c, 1, 2]
        if xp:
            assert np.array_equal(U[xp, 0:2, 2:18, 2:18], before[c, 16:18, 2:18, 2:18])
            assert np.array_equal(U[c, 18:20, 2:18, 2:18], before[xp, 2:4, 2:18, 2:18])
        yp = adj[c, 1, 4]
        if yp:
            assert np.array_equal(U[yp, 2:18, 0:2, 2:18], before[c, 2:18, 16:18, 2:18])
            assert np.array_equal(U[c, 2:18, 18:20, 2:18], before[yp, 2:18, 2:4, 2:18])
        # i-edge (j,k) = (1,1): mean of the two face neighbours
        assert np.array_equal(U[c, 2:18, 1, 1], 0.5 * (U[c, 2:18, 1, 2] + U[c, 2:18, 2, 1]))
        # k-edge (i,j) = (18,18)
        assert np.array_equal(U[c, 18, 18, 1:19], 0.5 * (U[c, 17, 18, 1:19] + U[c, 18, 17, 1:19]))
    # interior untouched
    assert np.array_equal(U[1:, 2:18, 2:18, 2:18], before[1:, 2:18, 2:18, 2:18])


def test_two_level_faces_follow_geometry(B, O):
    # On a two-level grid every small cube's ghost cell across a size change
    # takes the value of the big cube's cell that contains its centre, and
    # every big cube's ghost cell the mean of the 8 small cells inside it;
    # for a field linear in position that mean is the value at its centre.
    case = B.Case(B.case_params("ml_tiny", sphere_diameter=0.0, noise_amp=0.0))
    orc = O.OracleRun(case, nthreads=1)
    nc = case.ncube
    X, Y, Z = (case.array(a) for a in ("Xcnt", "Ycnt", "Zcnt"))
    size = case.array("cube_size")
    lin = np.array([0.3, 1.7, -0.9, 2.3])
    U = orc.state["U1_"]
    U[...] = 0.0
    for c in range(1, nc):
        f = (lin[0] + lin[1] * X[c][:, None, None] + lin[2] * Y[c][None, :, None]
             + lin[3] * Z[c][None, None, :])
        for v in range(5):
            U[c, 2:18, 2:18, 2:18, v] = (f * (v + 1))[2:18, 2:18, 2:18]
    orc.interface()
    adj = case.array("adj_number")
    csl = case.array("csl")
    checked_inj = checked_res = 0
    for c in range(1, nc):
        h = size[c] / 16
        for d in range(1, 7):
            nb = adj[c, 1:5, d]
            nb = nb[nb > 0]
            if len(nb) == 0 or csl[nb[0]] == csl[c]:
                continue
            ax, up = (d - 1) // 2, d % 2 == 0
            layers = (18, 19) if up else (0, 1)
            for l in layers:
                for p in range(2, 18):
                    for r in range(2, 18):
                        idx = [0, 0, 0]
                        t0, t1 = (1, 2) if ax == 0 else ((0, 2) if ax == 1 else (0, 1))
                        idx[ax], idx[t0], idx[t1] = l, p, r
                        centre = np.array([X[c][idx[0]], Y[c][idx[1]], Z[c][idx[2]]])
                        got = U[c, idx[0], idx[1], idx[2], 0]
                        if csl[c] < csl[nb[0]]:
                            # small cube: the big cell containing the centre
                            big = nb[0]
                            H = size[big] / 16
                            lo = np.array([case.array("Xcube")[big], case.array("Ycube")[big],
                                           case.array("Zcube")[big]])
                            cell = np.floor((centre - lo) / H).astype(int) + 2
                            assert got == U[big, cell[0], cell[1], cell[2], 0]
                            checked_inj += 1
                        elif p <= 17 and r <= 17:
                            # big cube: mean of the small cells inside, exact
                            # for the linear field up to rounding
                            expect = lin[0] + lin[1:] @ centre
                            assert abs(got - expect) <= 1e-12 * (1 + abs(expect))
                            checked_res += 1
    assert checked_inj > 0 and checked_res > 0
    # interior untouched
    for c in range(1, nc):
        assert np.all(np.isfinite(U[c, 2:18, 2:18, 2:18]))
