"""Pins three extension functions of the oracle by what the mathematics fixes,
each chosen so that a plausible slip fails:

* the minmod (Barth-Jespersen) limiter: the limited gradient is the unlimited
  one scaled by the LARGEST psi in [0, 1] that keeps the six Gauss-point
  reconstructions inside the cell/neighbour range -- psi = 1 where the linear
  reconstruction already stays inside, psi = 0 at a local extremum, and
  otherwise some Gauss point lands exactly on a bound (a halved or zeroed
  psi fails);
* the LAX (Rusanov) flux of the implicit method
  (src/methods/fvm_tvd_implicit.cpp:939-960): numerical viscosity alpha =
  max over both sides of |u.n| + c, seen in closed form on a stationary
  contact and on a pair whose faster side is not the one with the larger c
  (alpha = min, or max of c alone, fails);
* the Navier-Stokes heat conduction: with uniform p, zero velocity and a
  linear density the exact least-squares gradients give grad T =
  -p grad(rho) / (R rho^2) in every interior cell, and Fourier's law with
  k = mu Cp / Pr fixes each interior face's energy flux (mu Cp Pr, a lost
  sign or a dropped face mean fails).
"""
import math

import numpy as np

import oracle
from cfd2d import meshgen
from cfd2d.case import Case
from cfd2d.config import Boundary, Material, Region, SolverConfig, prim_to_cons

GAM = 1.4


def outlet_case(nx, ny, seed, **cfg):
    m = meshgen.rect_mesh(nx, ny, 1.0, 0.5, seed=seed)
    bnds = [Boundary(n, "BOUND_OUTLET") for n in m.edge_group_names]
    regions = [Region(n, 0, 0, 1, 1) for n in m.cell_group_names]
    return Case(m, SolverConfig(boundaries=bnds, regions=regions, max_newton=200, **cfg))


def neighbours(ex, nc):
    c1, c2 = ex["edge_int"][:, 2], ex["edge_int"][:, 3]
    nb = [[] for _ in range(nc)]
    for e in range(len(c1)):
        if c2[e] >= 0:
            nb[c1[e]].append(c2[e])
            nb[c2[e]].append(c1[e])
    return nb


def test_minmod_is_the_largest_bounded_scaling():
    nx, ny = 16, 8
    lim = oracle.OracleSolver(outlet_case(nx, ny, 11, limiter="MINMOD"))
    unl = oracle.OracleSolver(outlet_case(nx, ny, 11))
    ex = lim.geometry
    nc = lim.nc
    cen = ex["cell_geom"][:, 1:3]
    x, y = cen[:, 0], cen[:, 1]
    rho = 1.0 + 0.3 * np.sin(5.0 * x) * np.cos(7.0 * y) + 0.1 * x
    p = 1.0 + 0.2 * np.cos(6.0 * x + 3.0 * y)
    u = 0.2 * np.tanh(20.0 * (x - 0.5))  # a steep front
    v = 0.1 * np.sin(9.0 * y) + 0.05 * meshgen.uniform(3, 9, np.arange(nc))
    nb = neighbours(ex, nc)
    inner = np.array([len(nb[i]) == 3 for i in range(nc)])
    # an isolated maximum of rho and an isolated minimum of p
    imax = int(np.argmin((x - 0.3) ** 2 + (y - 0.25) ** 2))
    imin = int(np.argmin((x - 0.7) ** 2 + (y - 0.2) ** 2))
    rho[imax] = 3.0
    p[imin] = 0.2
    U = prim_to_cons(rho, u, v, p, GAM)
    lim.set_state(U)
    unl.set_state(U)
    Gl = lim.gradients().reshape(4, 2, nc)
    Gu = unl.gradients().reshape(4, 2, nc)
    # the oracle's primitives of U (p, u and v through the conserved variables)
    r_ = U[0]
    uu, vv = U[1] / r_, U[2] / r_
    phi = np.stack([r_, (U[3] / r_ - 0.5 * (uu * uu + vv * vv)) * r_ * (GAM - 1.0), uu, vv])
    c1 = ex["edge_int"][:, 2]
    gps = {i: [] for i in range(nc)}
    for e in range(len(c1)):
        for cell in (c1[e], ex["edge_int"][e, 3]):
            if cell >= 0:
                for g in range(2):
                    gps[cell].append(ex["edge_geom"][e, 5 + 2 * g: 7 + 2 * g] - cen[cell])
    n_one = n_mid = n_zero = 0
    for i in np.nonzero(inner)[0]:
        d = np.array(gps[i])  # (6, 2)
        assert d.shape == (6, 2)
        for q in range(4):
            lo = min(phi[q, i], phi[q, nb[i]].min())
            hi = max(phi[q, i], phi[q, nb[i]].max())
            gu, gl = Gu[q, :, i], Gl[q, :, i]
            scale = max(hi - lo, 1e-300)
            k = int(np.argmax(np.abs(gu)))
            if gu[k] == 0.0:
                assert np.all(gl == 0.0)
                continue
            psi = gl[k] / gu[k]
            # a scaling of the unlimited gradient, by psi in [0, 1]
            assert -1e-15 <= psi <= 1.0 + 1e-15, (i, q, psi)
            np.testing.assert_allclose(gl, psi * gu, rtol=1e-13, atol=1e-15 * np.abs(gu).max())
            Du = d @ gu
            inside = np.all(phi[q, i] + Du <= hi + 1e-14 * scale) and np.all(phi[q, i] + Du >= lo - 1e-14 * scale)
            val = phi[q, i] + d @ gl
            # every limited Gauss-point value inside the range
            assert np.all(val <= hi + 1e-12 * scale) and np.all(val >= lo - 1e-12 * scale), (i, q)
            if inside:
                assert psi == 1.0, (i, q, psi)
                n_one += 1
                continue
            # maximality: some Gauss point sits on the bound it moves towards
            bound = np.where(Du > 0, hi, lo)
            gap = np.where(Du != 0, np.abs(val - bound), np.inf).min()
            assert gap <= 1e-12 * scale, (i, q, psi, gap)
            if psi == 0.0:
                n_zero += 1
            else:
                n_mid += 1
    assert n_one > 50 and n_mid > 20 and n_zero >= 2, (n_one, n_mid, n_zero)
    # the two planted extrema: any Gauss point moving outward is already at the bound
    assert np.any(Gu[0, :, imax] != 0) and np.all(Gl[0, :, imax] == 0.0)
    assert np.any(Gu[1, :, imin] != 0) and np.all(Gl[1, :, imin] == 0.0)


def physical_flux(r, p, u, v, nx, ny):
    un = u * nx + v * ny
    E = p / (GAM - 1) + 0.5 * r * (u * u + v * v)
    return np.array([r * un, r * u * un + p * nx, r * v * un + p * ny, (E + p) * un])


def conserved(r, p, u, v):
    return np.array([r, r * u, r * v, p / (GAM - 1) + 0.5 * r * (u * u + v * v)])


def test_lax_stationary_contact_closed_form():
    """u.n = 0 and equal p on both sides: the mass flux is
    -1/2 max(c_L, c_R) (rho_R - rho_L), the normal momentum flux p, the
    tangential momentum and energy fluxes -1/2 alpha times their jumps."""
    rl, rr, p, vt = 1.0, 0.25, 0.7, 0.3
    F = oracle.flux(1, np.array([[rl, p, 0.0, vt, rr, p, 0.0, vt, 1.0, 0.0]]))[0]
    cl, cr = math.sqrt(GAM * p / rl), math.sqrt(GAM * p / rr)
    assert cr == 2.0 * cl  # the sides' speeds differ: max and min are told apart
    a = max(cl, cr)
    El, Er = p / (GAM - 1) + 0.5 * rl * vt * vt, p / (GAM - 1) + 0.5 * rr * vt * vt
    np.testing.assert_allclose(F, [-0.5 * a * (rr - rl), p, -0.5 * a * (rr - rl) * vt, -0.5 * a * (Er - El)],
                               rtol=1e-14, atol=1e-15)


def test_lax_viscosity_is_the_larger_wave_speed():
    """Rusanov: F = (F_L + F_R)/2 - alpha/2 (U_R - U_L), alpha = max(|un_L| + c_L,
    |un_R| + c_R).  Here the fast left side has the smaller sound speed, so
    alpha differs from max(c) + max|un| and from the right side's speed; the
    pair is rotated to a slanted normal (the |u.n|, not |u|, enters)."""
    th = 0.4
    nx, ny = math.cos(th), math.sin(th)
    L = (1.0, 1.0, 2.0 * nx - 0.2 * ny, 2.0 * ny + 0.2 * nx)  # u.n = 2, tangential 0.2
    R = (0.25, 1.0, 0.1 * nx + 0.5 * ny, 0.1 * ny - 0.5 * nx)  # u.n = 0.1, tangential -0.5
    F = oracle.flux(1, np.array([[*L, *R, nx, ny]]))[0]
    al = 2.0 + math.sqrt(GAM * 1.0 / 1.0)
    ar = 0.1 + math.sqrt(GAM * 1.0 / 0.25)
    assert al > ar and math.sqrt(GAM / 0.25) > math.sqrt(GAM / 1.0)
    ref = 0.5 * (physical_flux(*L, nx, ny) + physical_flux(*R, nx, ny)) - 0.5 * al * (conserved(*R) - conserved(*L))
    np.testing.assert_allclose(F, ref, rtol=1e-13, atol=1e-14)
    # swapping the sides and the normal negates it (the faster side is now on the right)
    Fs = oracle.flux(1, np.array([[*R, *L, -nx, -ny]]))[0]
    np.testing.assert_allclose(Fs, -ref, rtol=1e-13, atol=1e-14)


def test_heat_conduction_is_fourier_with_mu_cp_over_pr():
    """Uniform p, u = v = 0, linear rho: no inviscid mass or energy flux (a
    stationary contact at every Gauss point), no viscous stress, and each
    interior face's energy flux (both Gauss points times l/2) is
    -k grad(T)_face . n l with k = mu Cp / Pr and grad(T)_face the mean of
    the two cells' exact grad T = -p grad(rho) / (R rho^2)."""
    mu, pr, p0 = 3e-3, 0.72, 0.9
    grad_rho = np.array([0.4, -0.25])
    mat = Material(ML=mu)  # M = gR (R = 1), Cp = 3.5
    nx_, ny_ = 14, 8
    c = outlet_case(nx_, ny_, 5, gradient="LSQ", viscous=True, prandtl=pr, materials=[mat])
    s = oracle.OracleSolver(c)
    ex = s.geometry
    nc = s.nc
    cen = ex["cell_geom"][:, 1:3]
    rho = 1.0 + cen @ grad_rho
    s.set_state(prim_to_cons(rho, np.zeros(nc), np.zeros(nc), np.full(nc, p0), GAM))
    F = s.edge_fluxes()
    nb = neighbours(ex, nc)
    inner = np.array([len(nb[i]) == 3 for i in range(nc)])
    c1, c2 = ex["edge_int"][:, 2], ex["edge_int"][:, 3]
    k = mu * mat.Cp / pr
    gT = -p0 * grad_rho[None, :] / (rho[:, None] ** 2)  # R = 1
    checked = 0
    for e in range(s.ne):
        if c2[e] < 0 or not (inner[c1[e]] and inner[c2[e]]):
            continue
        n = ex["edge_geom"][e, :2]
        l = ex["edge_geom"][e, 2]
        gf = 0.5 * (gT[c1[e]] + gT[c2[e]])
        # the edge record is inviscid - viscous, summed over the Gauss points, times l/2
        np.testing.assert_allclose(F[e, 3], -k * (gf @ n) * l, rtol=1e-10, atol=1e-17)
        assert abs(F[e, 0]) <= 1e-15
        checked += 1
    assert checked > 100
    # and the conduction is what moves the energy: k grad T . n l is well above round-off
    assert np.abs(F[:, 3]).max() > 1e-4
