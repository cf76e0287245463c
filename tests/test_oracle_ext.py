"""Pins the oracle's extension functions (no reference code exists for them)
with closed forms, invariants and textbook results, each chosen so that a
plausible slip (a dropped term, a sign, a transposed index) fails."""
import math

import numpy as np
import pytest

import oracle
from cfd2d import case as cases
from cfd2d import meshgen
from cfd2d.case import Case
from cfd2d.config import Boundary, Material, Region, SolverConfig, prim_to_cons

GAM = 1.4


def linear_prims(c):
    x, y = c[:, 0], c[:, 1]
    rho = 1.0 + 0.1 * x + 0.05 * y
    u = 0.2 + 0.3 * x - 0.1 * y
    v = -0.1 + 0.05 * x + 0.2 * y
    p = 1.0 + 0.2 * x - 0.1 * y
    grads = np.array([[0.1, 0.05], [0.2, -0.1], [0.3, -0.1], [0.05, 0.2]])  # r, p, u, v
    return rho, u, v, p, grads


def rect_case(nx=12, ny=6, **cfg):
    m = meshgen.rect_mesh(nx, ny, 1.0, 0.5, seed=3)
    bnds = [Boundary(n, "BOUND_OUTLET") for n in m.edge_group_names]
    regions = [Region(n, 0, 0, 1, 1) for n in m.cell_group_names]
    return Case(m, SolverConfig(boundaries=bnds, regions=regions, max_newton=200, **cfg))


def interior_cells(s):
    ce = s.geometry["cell_edges"]
    c2 = s.geometry["edge_int"][:, 3]
    return np.all(c2[ce] >= 0, axis=1)


def test_lsq_exact_for_linear_fields():
    case = rect_case(gradient="LSQ")
    s = oracle.OracleSolver(case)
    c = s.geometry["cell_geom"][:, 1:3]
    rho, u, v, p, grads = linear_prims(c)
    s.set_state(prim_to_cons(rho, u, v, p, GAM))
    G = s.gradients().reshape(4, 2, -1)
    inner = interior_cells(s)
    assert inner.sum() > 50
    for q in range(4):
        for d in range(2):
            np.testing.assert_allclose(G[q, d, inner], grads[q, d], rtol=0, atol=1e-12)


def test_periodic_seam_offsets_and_uniform_state():
    """Every joined seam edge sees its c2 centroid, shifted, next to c1 (the
    shift has the right sign and direction), and a uniform state has zero
    LSQ gradients and zero residual in every cell, seam cells included."""
    n = 16
    case = cases.periodic_pulse(n, seed=4, amp=0.0, noise=0.0, viscous=False)
    s = oracle.OracleSolver(case)
    ex = s.geometry
    c = ex["cell_geom"][:, 1:3]
    c1, c2 = ex["edge_int"][:, 2], ex["edge_int"][:, 3]
    d = c[c2] + ex["edge_shift"] - c[c1]
    assert (np.abs(ex["edge_shift"]).sum(axis=1) > 0).sum() == 2 * n  # n left-right + n bottom-top pairs
    assert np.linalg.norm(d, axis=1).max() < 1.5 / n
    assert np.abs(s.gradients()).max() < 1e-13
    assert np.abs(s.residual()).max() < 1e-14


def test_minmod_bounds():
    """Limited reconstruction at every Gauss point stays within the min/max
    of the cell and its face neighbours (the Barth-Jespersen property)."""
    case = cases.shock_tube(30, 5, seed=8, mild=False, limiter="MINMOD", max_newton=200)
    s = oracle.OracleSolver(case)
    s.run(5)
    U = s.state()
    G = s.gradients().reshape(4, 2, -1)
    ex = s.geometry
    r = U[0]
    uu = U[1] / r
    vv = U[2] / r
    p = (U[3] / r - 0.5 * (uu * uu + vv * vv)) * r * (GAM - 1.0)
    phi = np.stack([r, p, uu, vv])
    cent = ex["cell_geom"][:, 1:3]
    c1, c2 = ex["edge_int"][:, 2], ex["edge_int"][:, 3]
    nb = [[] for _ in range(s.nc)]
    for e in range(s.ne):
        if c2[e] >= 0:
            nb[c1[e]].append(c2[e])
            nb[c2[e]].append(c1[e])
    worst = 0.0
    for e in range(s.ne):
        for side, cell in ((0, c1[e]), (1, c2[e])):
            if cell < 0 or len(nb[cell]) < 3:
                continue
            lo = np.minimum(phi[:, cell], phi[:, nb[cell]].min(axis=1))
            hi = np.maximum(phi[:, cell], phi[:, nb[cell]].max(axis=1))
            for g in range(2):
                gp = ex["edge_geom"][e, 5 + 2 * g: 7 + 2 * g]
                d = gp - cent[cell]
                val = phi[:, cell] + G[:, 0, cell] * d[0] + G[:, 1, cell] * d[1]
                worst = max(worst, float(np.max(lo - val)), float(np.max(val - hi)))
    assert worst <= 1e-13, worst


def random_states(n, seed=1):
    i = np.arange(n)
    u = [0.5 * (meshgen.uniform(seed, 60 + k, i) + 1.0) for k in range(10)]
    ang = 2 * math.pi * u[8]
    return np.stack([0.2 + u[0], 0.2 + u[1], u[2] - 0.5, u[3] - 0.5,
                     0.2 + u[4], 0.2 + u[5], u[6] - 0.5, u[7] - 0.5, np.cos(ang), np.sin(ang)], axis=1)


def physical_flux(r, p, u, v, nx, ny):
    un = u * nx + v * ny
    E = p / (GAM - 1) + 0.5 * r * (u * u + v * v)
    return np.stack([r * un, r * u * un + p * nx, r * v * un + p * ny, (E + p) * un], axis=-1)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_flux_consistency(kind):
    """F(U, U) = physical normal flux, for every flux."""
    rows = random_states(500)
    rows[:, 4:8] = rows[:, 0:4]
    F = oracle.flux(kind, rows)
    ref = physical_flux(rows[:, 0], rows[:, 1], rows[:, 2], rows[:, 3], rows[:, 8], rows[:, 9])
    np.testing.assert_allclose(F, ref, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_flux_rotation_and_mirror(kind):
    """Rotating states and normal together rotates the momentum flux; swapping
    sides and flipping the normal negates the flux (conservation)."""
    rows = random_states(300, seed=2)
    F = oracle.flux(kind, rows)
    th = 0.7
    c, s_ = math.cos(th), math.sin(th)
    rot = rows.copy()
    for a, b in ((2, 3), (6, 7), (8, 9)):
        rot[:, a] = c * rows[:, a] - s_ * rows[:, b]
        rot[:, b] = s_ * rows[:, a] + c * rows[:, b]
    Fr = oracle.flux(kind, rot)
    np.testing.assert_allclose(Fr[:, 0], F[:, 0], rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(Fr[:, 3], F[:, 3], rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(Fr[:, 1], c * F[:, 1] - s_ * F[:, 2], rtol=1e-9, atol=1e-11)
    sw = np.concatenate([rows[:, 4:8], rows[:, 0:4], -rows[:, 8:10]], axis=1)
    Fs = oracle.flux(kind, sw)
    np.testing.assert_allclose(Fs, -F, rtol=1e-9, atol=1e-11)


def test_roe_single_moving_shock():
    """Data joined by one shock moving at +0.5 (the stationary M1 = 2 shock
    seen from a frame moving at -0.5): Roe's linearisation puts that shock
    on its first characteristic, lambda_1 = 0.5 > 0.1 c (the entropy fix is
    inactive), so the flux is F(U_L) (Roe 1981; Toro 11.2); the exact
    solution sampled at x/t = 0 is also the left state."""
    M1, r1, p1 = 2.0, 1.0, 1.0
    c1 = math.sqrt(GAM * p1 / r1)
    u1 = M1 * c1
    r2 = r1 * (GAM + 1) * M1 ** 2 / ((GAM - 1) * M1 ** 2 + 2)
    p2 = p1 * (2 * GAM * M1 ** 2 - (GAM - 1)) / (GAM + 1)
    u2 = u1 * r1 / r2
    w = 0.5
    row = np.array([[r1, p1, u1 + w, 0.1, r2, p2, u2 + w, 0.1, 1.0, 0.0]])
    ref = physical_flux(r1, p1, u1 + w, 0.1, 1.0, 0.0)
    np.testing.assert_allclose(oracle.flux(3, row)[0], ref, rtol=1e-12)
    np.testing.assert_allclose(oracle.flux(0, row)[0], ref, rtol=1e-12)


def test_supersonic_upwinding():
    """Both sides supersonic in +n: Godunov, HLLC and Roe return F(U_L)."""
    row = np.array([[1.0, 1.0, 3.0, 0.2, 0.5, 0.6, 2.8, -0.1, 1.0, 0.0]])
    ref = physical_flux(1.0, 1.0, 3.0, 0.2, 1.0, 0.0)
    for kind in (0, 2, 3):
        np.testing.assert_allclose(oracle.flux(kind, row)[0], ref, rtol=1e-12)


def test_hllc_resolves_a_contact():
    """A stationary contact (equal p, zero normal velocity, density jump) is
    exact for HLLC: zero mass/energy flux, momentum flux = p n."""
    row = np.array([[1.0, 0.7, 0.0, 0.3, 0.25, 0.7, 0.0, 0.3, 1.0, 0.0]])
    F = oracle.flux(2, row)[0]
    np.testing.assert_allclose(F, [0.0, 0.7, 0.0, 0.0], atol=1e-14)


def test_viscous_shear_heating():
    """Linear shear u = a y, v = 0 at uniform rho, p: with exact (LSQ) face
    gradients the viscous momentum residual vanishes and the viscous energy
    residual of an interior cell equals mu a^2 S (viscous dissipation)."""
    a, mu = 0.3, 2e-3

    def resid(viscous):
        m = meshgen.rect_mesh(16, 8, 1.0, 0.5, seed=5)
        bnds = [Boundary(n, "BOUND_OUTLET") for n in m.edge_group_names]
        mat = Material(ML=mu)
        c = Case(m, SolverConfig(materials=[mat], boundaries=bnds, gradient="LSQ", viscous=viscous,
                                 flux="LAX", regions=[Region(n, 0, 0, 1, 1) for n in m.cell_group_names]))
        s = oracle.OracleSolver(c)
        cc = s.geometry["cell_geom"][:, 1:3]
        n = s.nc
        s.set_state(prim_to_cons(np.ones(n), a * cc[:, 1], np.zeros(n), np.ones(n), GAM))
        return s, s.residual()

    s, Rv = resid(True)
    _, Ri = resid(False)
    dR = Ri - Rv  # the viscous part enters with a minus sign
    inner = interior_cells(s)
    # only cells whose neighbours also have exact gradients
    ce = s.geometry["cell_edges"]
    c1, c2 = s.geometry["edge_int"][:, 2], s.geometry["edge_int"][:, 3]
    ok = inner.copy()
    for i in np.nonzero(inner)[0]:
        for e in ce[i]:
            j = c2[e] if c1[e] == i else c1[e]
            ok[i] &= inner[j]
    S = s.geometry["cell_geom"][:, 0]
    assert ok.sum() > 30
    np.testing.assert_allclose(dR[1:3, ok], 0.0, atol=1e-15)
    np.testing.assert_allclose(dR[3, ok], -mu * a * a * S[ok], rtol=1e-9)


def rk_error_ratio(stages):
    """Temporal self-convergence on a smooth periodic flow (fixed mesh)."""
    case = cases.periodic_pulse(12, seed=9, noise=0.0, viscous=False, flux="LAX", rk_stages=stages,
                                gradient="LSQ")
    T = 0.02
    out = []
    for k in (1, 2, 4):
        s = oracle.OracleSolver(case)
        n = 8 * k
        s.set_tau(T / n)
        s.run(n)
        out.append(s.state())
    e1 = np.abs(out[0] - out[1]).max()
    e2 = np.abs(out[1] - out[2]).max()
    return e1 / e2


def test_ssp_rk3_third_order():
    assert 7.0 < rk_error_ratio(3) < 9.0


def test_heun_rk2_second_order():
    assert 3.5 < rk_error_ratio(2) < 4.5


def test_periodic_conservation_and_freestream():
    case = cases.periodic_pulse(20, seed=6)
    s = oracle.OracleSolver(case)
    S = s.geometry["cell_geom"][:, 0]
    t0 = (s.state() * S).sum(axis=1)
    s.run(30)
    t1 = (s.state() * S).sum(axis=1)
    assert np.all(np.abs(t1 - t0) <= 1e-14 * np.abs(t0).max())
    fs = cases.periodic_pulse(20, seed=6, amp=0.0, noise=0.0)
    f = oracle.OracleSolver(fs)
    U0 = f.state()
    f.run(10)
    assert np.abs(f.state() - U0).max() <= 1e-13


def test_spectral_time_step_closed_form():
    """Uniform state: lambda = |V| + c on every face, so
    tau = min_i 2 CFL S_i / ((|V| + c) * perimeter_i)."""
    case = cases.periodic_pulse(10, seed=2, amp=0.0, noise=0.0, cfl_mode="SPECTRAL", cfl=0.4)
    s = oracle.OracleSolver(case)
    ex = s.geometry
    S = ex["cell_geom"][:, 0]
    per = ex["edge_geom"][:, 2][ex["cell_edges"]].sum(axis=1)
    lam = 0.5 + 1.0  # |V| = Mach 0.5 with c = 1
    np.testing.assert_allclose(s.tau()[0], (2 * 0.4 * S / (lam * per)).min(), rtol=1e-13)


def test_pressure_outlet_ghost():
    """Subsonic outflow: at each Gauss point of an outlet face the ghost is
    the reconstructed interior state with the prescribed pressure; the face
    flux is the sum of the two Godunov fluxes times l/2."""
    m = meshgen.rect_mesh(6, 3, 1.0, 0.5, seed=2)
    bnds = [Boundary("edge_right", "BOUND_PRESSURE_OUTLET", P=0.8), Boundary("edge_left", "BOUND_OUTLET"),
            Boundary("edge_bottom", "BOUND_WALL_SLIP"), Boundary("edge_top", "BOUND_WALL_SLIP")]
    c = Case(m, SolverConfig(boundaries=bnds, regions=[Region(n, 0.3, 0.0, 1.0, 1.0) for n in m.cell_group_names],
                             max_newton=200))
    s = oracle.OracleSolver(c)
    F = s.edge_fluxes()
    G = s.gradients().reshape(4, 2, -1)
    ex = s.geometry
    right = np.nonzero(ex["edge_group"] == m.group_index("edge_right"))[0]
    assert len(right) == 3
    for e in right:
        nx, ny, l = ex["edge_geom"][e, :3]
        c1 = ex["edge_int"][e, 2]
        cen = ex["cell_geom"][c1, 1:3]
        acc = np.zeros(4)
        for g in range(2):
            d = ex["edge_geom"][e, 5 + 2 * g: 7 + 2 * g] - cen
            r, p, u, v = np.array([1.0, 1.0, 0.3, 0.0]) + G[:, 0, c1] * d[0] + G[:, 1, c1] * d[1]
            row = np.array([[r, p, u, v, r, 0.8, u, v, nx, ny]])
            acc += oracle.flux(0, row)[0]
        np.testing.assert_allclose(F[e], acc * (l * 0.5), rtol=1e-12)


def test_uniform_tangential_flow_is_preserved():
    """Reference-exact mode (Green-Gauss, exact Godunov, Heun RK2) with slip
    walls and outlets: a uniform flow along the walls has zero residual and
    zero mass flux through the walls."""
    case = cases.shock_tube(20, 4, seed=3, mild=True)
    for r in case.config.regions:
        r.Vx, r.T, r.P = 0.4, 1.0, 1.0
    s = oracle.OracleSolver(case)
    R = s.residual()
    assert np.abs(R).max() < 1e-14
    F = s.edge_fluxes()
    ex = s.geometry
    walls = np.isin(ex["edge_group"], [case.mesh.group_index("edge_bottom"), case.mesh.group_index("edge_top")])
    assert walls.sum() == 40
    assert np.abs(F[walls][:, [0, 3]]).max() < 1e-16


