"""The tangent-linear model (oracle/tangent.py) that the GPU adjoint is
checked against: its value part must equal the C oracle, and its tangents
must match central finite differences of the oracle's own forward map."""
import numpy as np
import pytest

import oracle
import tangent
from cfd2d import case as cases


def bump(flux, gradient, rk, ny=6):
    return cases.bump_channel(ny, height=0.06, mach=0.8, seed=3, flux=flux, gradient=gradient, limiter="NONE",
                              rk_stages=rk)


def pulse(flux, gradient, rk, n=12):
    return cases.periodic_pulse(n, seed=2, flux=flux, gradient=gradient, rk_stages=rk, viscous=False,
                                cfl_mode="REFERENCE")


def bump_mask(s):
    g = s.case.mesh.group_index("bump")
    return (s.geometry["edge_group"] == g).astype(np.int32)


def directions(s, seed=5):
    rng = np.random.default_rng(seed)
    U0 = s.case.initial_state()
    v = rng.standard_normal(U0.shape) * np.abs(U0) * 1e-2
    m = tangent.Model(s)
    w = np.zeros_like(m.bcpar0)
    w[0] = [1.0, 0.5, 0.2, 0.3]
    if len(w) > 1:
        w[1] = [0.7, 0.0, 0.0, 0.0]
    return U0, v, w


@pytest.mark.parametrize("make,flux,gradient,rk", [
    (bump, "LAX", "GREEN_GAUSS", 2), (bump, "HLLC", "LSQ", 3), (bump, "LAX", "LSQ", 2),
    (pulse, "HLLC", "GREEN_GAUSS", 3), (bump, "GODUNOV", "GREEN_GAUSS", 2), (pulse, "GODUNOV", "LSQ", 3)])
def test_model_forward_equals_oracle(make, flux, gradient, rk):
    c = make(flux, gradient, rk)
    s = oracle.OracleSolver(c)
    tau, _ = s.tau()
    U0 = c.initial_state()
    U = tangent.forward(s, U0, 6, tau)
    s.run(6)
    ref = s.state()
    scale = np.abs(ref).max(axis=1, keepdims=True)
    assert np.max(np.abs(U - ref) / scale) <= 1e-14


@pytest.mark.parametrize("flux,gradient,rk,tol", [("LAX", "GREEN_GAUSS", 2, 5e-5), ("HLLC", "LSQ", 3, 2e-6),
                                                  ("GODUNOV", "GREEN_GAUSS", 2, 2e-6), ("GODUNOV", "LSQ", 3, 1e-5)])
def test_tangent_matches_finite_differences(flux, gradient, rk, tol):
    """dJ[v, w] against (J(x + h d) - J(x - h d)) / 2h with the oracle's forward
    map; LAX's max() wave speed switches inside the stencil, which bounds the
    finite-difference accuracy there."""
    c = bump(flux, gradient, rk)
    s = oracle.OracleSolver(c)
    tau, _ = s.tau()
    mask = bump_mask(s)
    U0, v, w = directions(s)
    n = 20
    J, dJ = tangent.tangent(s, U0, v, w, n, tau, mask)
    m = tangent.Model(s)

    def Jof(h):
        m.bcpar0 = tangent.Model(s).bcpar0 + h * w
        par = [m.bcpar0[:, k].copy() for k in range(4)]
        U = m.run([(U0 + h * v)[q].copy() for q in range(4)], par, n, tau)
        return float(m.objective(U, mask))

    h = 1e-6
    fd = (Jof(h) - Jof(-h)) / (2 * h)
    assert abs(J - Jof(0.0)) <= 1e-14 * abs(J)
    assert abs(dJ - fd) <= tol * abs(fd), (dJ, fd)


def test_tangent_objective_is_the_wall_force():
    """J equals the y pressure force that cfd2d_wall_force / the oracle's
    calcLiftForce restatement reports for the same edges."""
    c = bump("HLLC", "GREEN_GAUSS", 2)
    s = oracle.OracleSolver(c)
    tau, _ = s.tau()
    mask = bump_mask(s)
    J, _ = tangent.tangent(s, c.initial_state(), np.zeros((4, s.nc)), np.zeros((5, 4)), 3, tau, mask)
    s.run(3)
    _, fy = s.wall_force(mask)
    assert abs(J - fy) <= 1e-13 * abs(fy)


def bump_nodes(c):
    m = c.mesh
    return np.unique(m.bedges[m.bedge_group == m.group_index("bump")])


@pytest.mark.parametrize("flux,gradient,rk,tol", [("LAX", "GREEN_GAUSS", 2, 5e-5), ("HLLC", "LSQ", 3, 2e-6),
                                                  ("HLLC", "GREEN_GAUSS", 2, 2e-6), ("GODUNOV", "GREEN_GAUSS", 2, 2e-6)])
def test_geometry_tangent_matches_finite_differences(flux, gradient, rk, tol):
    """dJ[dxy] of the tangent model with the geometry rebuilt as Duals from the
    nodes, against central differences of the C oracle run on meshes with
    the nodes moved (bump nodes in y, a few interior nodes in x and y; the
    time step held): pins the model's geometric derivatives (normals,
    lengths, Gauss points, centroids, areas, weights, the objective's l n_y)."""
    import dataclasses
    c = bump(flux, gradient, rk)
    s = oracle.OracleSolver(c)
    tau, _ = s.tau()
    mask = bump_mask(s)
    n = 15
    xy = c.mesh.xy
    rng = np.random.default_rng(3)
    dxy = np.zeros_like(xy)
    bn = bump_nodes(c)
    dxy[bn, 1] = rng.standard_normal(len(bn)) * 0.01
    inner = rng.choice(np.setdiff1d(np.arange(len(xy)), np.unique(c.mesh.bedges)), 12, replace=False)
    dxy[inner] = rng.standard_normal((12, 2)) * 0.01
    J, dJ = tangent.tangent_geometry(s, xy, dxy, n, tau, mask)

    def Jof(h):
        cc = dataclasses.replace(c, mesh=dataclasses.replace(c.mesh, xy=xy + h * dxy))
        o = oracle.OracleSolver(cc)
        o.set_tau(tau)
        o.run(n)
        return o.wall_force(mask)[1]

    J0 = Jof(0.0)
    assert abs(J - J0) <= 1e-13 * abs(J0)
    h = 1e-5
    fd = (Jof(h) - Jof(-h)) / (2 * h)
    assert abs(dJ) > 1e-6 * abs(J)
    assert abs(dJ - fd) <= tol * abs(fd), (dJ, fd)
