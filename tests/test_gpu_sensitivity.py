"""Discrete adjoint on the GPU (cfd2d_adjoint) against the tangent-linear
model of the oracle (oracle/tangent.py) and against finite differences of the
GPU's own forward map.  The exact Godunov flux (the reference's) is
differentiated through its Newton iterations as evaluated (forward mode on
the device, a row-vectorised dual restatement in the model).

Both sides differentiate the same discrete map exactly, so they agree to
round-off: <dJ/dU0, v> + <dJ/dpar, w> within 1e-9 relative of the tangent,
each boundary-parameter derivative within 1e-9 of the largest one, J within
1e-12.  At 21,600 cells the check is against central differences (h = 1e-6,
relative 1e-6; measured 3e-8 with HLLC), which holds where no max()/min()
switch sits near a tie.
"""
import numpy as np
import pytest

import oracle
import tangent
from cfd2d import case as cases
from cfd2d import native

pytestmark = pytest.mark.gpu


def bump(flux, gradient, rk, ny=6):
    return cases.bump_channel(ny, height=0.06, mach=0.8, seed=3, flux=flux, gradient=gradient, limiter="NONE",
                              rk_stages=rk)


def pulse(flux, gradient, rk, n=12):
    return cases.periodic_pulse(n, seed=2, flux=flux, gradient=gradient, rk_stages=rk, viscous=False,
                                cfl_mode="REFERENCE")


def objective_mask(s, kind):
    if kind == "bump":
        # bump and bottom wall: the bottom wall meets the inlet and the outlet,
        # so every boundary parameter acts on J within the run
        eg = s.geometry["edge_group"]
        m = s.case.mesh
        return ((eg == m.group_index("bump")) | (eg == m.group_index("wall_bottom"))).astype(np.int32)
    c1 = s.geometry["edge_int"][:, 2]
    return (c1 < s.nc // 4).astype(np.int32)  # a band of interior edges


CONFIGS = [("bump", "LAX", "GREEN_GAUSS", 2), ("bump", "HLLC", "LSQ", 3), ("bump", "LAX", "LSQ", 3),
           ("bump", "HLLC", "GREEN_GAUSS", 2), ("pulse", "HLLC", "LSQ", 2), ("pulse", "LAX", "GREEN_GAUSS", 3),
           ("bump", "GODUNOV", "GREEN_GAUSS", 2), ("bump", "GODUNOV", "LSQ", 3), ("pulse", "GODUNOV", "LSQ", 3)]


@pytest.mark.parametrize("kind,flux,gradient,rk", CONFIGS)
def test_adjoint_matches_tangent(kind, flux, gradient, rk):
    make = bump if kind == "bump" else pulse
    c = make(flux, gradient, rk)
    s = oracle.OracleSolver(c)
    tau, _ = s.tau()
    mask = objective_mask(s, kind)
    n = 12
    g = native.GpuSolver(c)
    g.set_tau(tau)
    J, gU, gb = g.adjoint(n, mask, ckpt_every=5)
    g.close()
    rng = np.random.default_rng(11)
    U0 = c.initial_state()
    v = rng.standard_normal(U0.shape) * np.abs(U0) * 1e-2
    m = tangent.Model(s)
    w = rng.standard_normal(m.bcpar0.shape) * np.abs(m.bcpar0) * 1e-2
    Jm, dJ = tangent.tangent(s, U0, v, w, n, tau, mask)
    assert abs(J - Jm) <= 1e-12 * abs(Jm)
    ad = float(np.sum(gU * v)) + float(np.sum(gb * w))
    assert abs(ad - dJ) <= 1e-9 * abs(dJ), (ad, dJ)
    if kind == "bump":
        # each inlet parameter (u, v, T, p) and the outlet pressure separately
        zero = np.zeros_like(U0)
        big = np.max(np.abs(gb))
        assert big > 1e-8 * abs(J)
        for bc, k in [(0, 0), (0, 1), (0, 2), (0, 3), (1, 0)]:
            e = np.zeros_like(m.bcpar0)
            e[bc, k] = 1.0
            _, d = tangent.tangent(s, U0, zero, e, n, tau, mask)
            assert abs(gb[bc, k] - d) <= 1e-9 * big, (bc, k, gb[bc, k], d)


def test_adjoint_checkpoint_interval_and_final_state():
    """The reverse sweep recomputes each segment deterministically, so the
    gradient is bitwise the same for any checkpoint interval, and the context
    ends at the same state as plain steps."""
    c = bump("HLLC", "LSQ", 3, ny=10)
    n = 13
    g = native.GpuSolver(c)
    mask = np.zeros(g.ne, np.int32)
    s = oracle.OracleSolver(c)
    mask[:] = (s.geometry["edge_group"] == c.mesh.group_index("bump"))
    res = []
    for ck in (1, 4, 13, 50, 0):
        g.set_state(c.initial_state())
        res.append(g.adjoint(n, mask, ckpt_every=ck))
        res[-1] = res[-1] + (g.state(),)
    ref = native.GpuSolver(c)
    ref.run(n)
    Uref = ref.state()
    ref.close()
    g.close()
    for J, gU, gb, U in res[1:]:
        assert J == res[0][0]
        assert np.array_equal(gU, res[0][1])
        assert np.array_equal(gb, res[0][2])
        assert np.array_equal(U, res[0][3])
    assert np.array_equal(res[0][3], Uref)


def test_adjoint_finite_differences_21k_cells():
    c = bump("HLLC", "LSQ", 3, ny=60)
    s_or = oracle.OracleSolver(c)
    mask = (s_or.geometry["edge_group"] == c.mesh.group_index("bump")).astype(np.int32)
    del s_or
    g = native.GpuSolver(c)
    n = 30
    U0 = c.initial_state()
    J, gU, gb = g.adjoint(n, mask, ckpt_every=8)
    rng = np.random.default_rng(3)
    v = rng.standard_normal(U0.shape) * np.abs(U0) * 1e-2

    def Jof(h):
        g.set_state(U0 + h * v)
        g.step(n)
        g.sync()
        return g.wall_force(mask)[1]

    h = 1e-6
    fd = (Jof(h) - Jof(-h)) / (2 * h)
    assert abs(Jof(0.0) - J) <= 1e-13 * abs(J)
    ad = float(np.sum(gU * v))
    g.close()
    print("adjoint %.15e fd %.15e rel %.3e" % (ad, fd, abs(ad - fd) / abs(fd)))
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)


@pytest.mark.parametrize("over", [dict(flux="ROE"), dict(limiter="MINMOD"), dict(u_max=1e-3)])
def test_adjoint_rejects_unsupported(over):
    """Configurations without a reverse are refused (code 5): the Roe flux,
    the limiter, and a run in which a cell leaves the bounds."""
    cfg = dict(flux="LAX", gradient="GREEN_GAUSS", limiter="NONE", rk_stages=2)
    cfg.update(over)
    c = cases.bump_channel(6, height=0.06, mach=0.8, seed=3, **cfg)
    g = native.GpuSolver(c)
    with pytest.raises(native.Cfd2dError) as ei:
        g.adjoint(2, np.zeros(g.ne, np.int32))
    assert ei.value.code == 5
    with pytest.raises(native.Cfd2dError) as ei:
        g.adjoint(0, np.zeros(g.ne, np.int32))
    assert ei.value.code == 1
    g.close()


def test_mach_and_angle_gradients_match_finite_differences():
    """dJ/dMach and dJ/dangle through the inlet parameters and the
    free-stream initial state (cfd2d.sensitivity) against central differences
    of cases rebuilt at Mach +- h and angle +- h, time step held fixed."""
    from cfd2d import sensitivity

    def make(mach, ang):
        return cases.bump_channel(16, height=0.06, mach=mach, angle_deg=ang, seed=3, flux="HLLC", gradient="GREEN_GAUSS",
                                  limiter="NONE", rk_stages=2)

    M, A, n = 0.8, 2.0, 30
    c = make(M, A)
    g = native.GpuSolver(c)
    tau, _ = g.tau()
    mask = g.edge_mask("bump") | g.edge_mask("wall_bottom")
    J, gU, gb = g.adjoint(n, mask)
    g.close()
    dM, dA = sensitivity.mach_angle_gradient(c, gU, gb)

    def Jat(mach, ang):
        s = native.GpuSolver(make(mach, ang))
        s.set_tau(tau)
        s.run(n)
        j = s.wall_force(mask)[1]
        s.close()
        return j

    fdM = (Jat(M + 1e-6, A) - Jat(M - 1e-6, A)) / 2e-6
    fdA = (Jat(M, A + 1e-5) - Jat(M, A - 1e-5)) / 2e-5
    print("dJ/dMach %.12e fd %.12e; dJ/dangle %.12e fd %.12e" % (dM, fdM, dA, fdA))
    assert abs(Jat(M, A) - J) <= 1e-13 * abs(J)
    assert abs(dM - fdM) <= 1e-6 * abs(fdM)
    assert abs(dA - fdA) <= 1e-5 * abs(fdA)


@pytest.mark.parametrize("flux,gradient,rk", [("LAX", "GREEN_GAUSS", 2), ("HLLC", "LSQ", 3)])
def test_batched_adjoint_equals_individual(flux, gradient, rk):
    """One reverse sweep over concatenated cases gives each case's own
    gradients: dJ/dU0 bit for bit (per-cell work is identical), J and the
    boundary-parameter sums to round-off (the per-object reduction splits the
    boundary edges over threads differently)."""
    from cfd2d import batch
    specs = [(5, 0.7, 1.0, 0.04), (7, 1.3, -2.0, 0.08), (9, 0.5, 4.0, 0.06)]
    cs = [cases.bump_channel(ny, height=h, mach=m, angle_deg=a, seed=ny, flux=flux, gradient=gradient,
                             limiter="NONE", rk_stages=rk) for ny, m, a, h in specs]
    n = 12
    b = batch.BatchSolver(cs)
    mask = b.edge_mask("bump") | b.edge_mask("wall_bottom")
    res = b.adjoint(n, mask, ckpt_every=5)
    b.close()
    for c, (Jb, gUb, gbb) in zip(cs, res):
        s = native.GpuSolver(c)
        m = s.edge_mask("bump") | s.edge_mask("wall_bottom")
        J, gU, gb = s.adjoint(n, m, ckpt_every=5)
        s.close()
        assert abs(Jb - J) <= 1e-13 * abs(J)
        assert np.array_equal(gUb, gU)
        assert np.max(np.abs(gbb - gb)) <= 1e-13 * np.max(np.abs(gb))


def _bump_nodes(c):
    m = c.mesh
    return np.unique(m.bedges[m.bedge_group == m.group_index("bump")])


@pytest.mark.parametrize("flux,gradient,rk", [("LAX", "GREEN_GAUSS", 2), ("HLLC", "LSQ", 3), ("LAX", "LSQ", 2),
                                              ("HLLC", "GREEN_GAUSS", 3), ("GODUNOV", "GREEN_GAUSS", 2),
                                              ("GODUNOV", "LSQ", 3)])
def test_geometry_adjoint_matches_tangent(flux, gradient, rk):
    """cfd2d_adjoint_geometry: <dJ/dxy, dxy> equals the tangent model's dJ[dxy]
    with the geometry rebuilt from the nodes (every node moved) to 1e-9, and
    its dJ/dU0, dJ/dpar equal cfd2d_adjoint's."""
    c = bump(flux, gradient, rk)
    s = oracle.OracleSolver(c)
    tau, _ = s.tau()
    mask = objective_mask(s, "bump")
    n = 12
    g = native.GpuSolver(c)
    g.set_tau(tau)
    J, gU, gb, gx = g.adjoint_geometry(n, mask, ckpt_every=5)
    g.set_state(c.initial_state())
    J2, gU2, gb2 = g.adjoint(n, mask, ckpt_every=5)
    g.close()
    assert J == J2
    assert np.max(np.abs(gU - gU2)) <= 1e-13 * np.max(np.abs(gU2))
    assert np.max(np.abs(gb - gb2)) <= 1e-13 * np.max(np.abs(gb2))
    rng = np.random.default_rng(5)
    dxy = rng.standard_normal(c.mesh.xy.shape) * 1e-2
    Jm, dJ = tangent.tangent_geometry(s, c.mesh.xy, dxy, n, tau, mask)
    ad = float(np.sum(gx * dxy))
    print("geometry: adjoint %.15e tangent %.15e rel %.2e" % (ad, dJ, abs(ad - dJ) / abs(dJ)))
    assert abs(J - Jm) <= 1e-12 * abs(Jm)
    assert abs(ad - dJ) <= 1e-9 * abs(dJ), (ad, dJ)


def test_geometry_adjoint_bump_nodes_finite_differences():
    """dJ/d(bump-node y) against central differences of the GPU forward map
    on meshes with one bump node moved (h = 1e-6, time step held), for nodes
    along the bump, and one interior node in x and y."""
    import dataclasses
    c = bump("HLLC", "GREEN_GAUSS", 2, ny=16)
    g = native.GpuSolver(c)
    tau, _ = g.tau()
    mask = g.edge_mask("bump")
    n = 30
    J, _, _, gx = g.adjoint_geometry(n, mask, ckpt_every=10)
    g.close()
    bn = _bump_nodes(c)
    xs = c.mesh.xy[bn, 0]
    picks = [(int(bn[np.argmin(np.abs(xs - x))]), 1) for x in (1.2, 1.5, 1.8)]
    inner = int(np.argmin(np.sum((c.mesh.xy - [1.4, 0.2]) ** 2, axis=1)))
    picks += [(inner, 0), (inner, 1)]

    def Jat(v, ax, h):
        xy = c.mesh.xy.copy()
        xy[v, ax] += h
        s = native.GpuSolver(dataclasses.replace(c, mesh=dataclasses.replace(c.mesh, xy=xy)))
        s.set_tau(tau)
        s.run(n)
        j = s.wall_force(mask)[1]
        s.close()
        return j

    assert abs(Jat(0, 0, 0.0) - J) <= 1e-13 * abs(J)
    h = 1e-6
    for v, ax in picks:
        fd = (Jat(v, ax, h) - Jat(v, ax, -h)) / (2 * h)
        print("node %d axis %d: adjoint %.12e fd %.12e" % (v, ax, gx[v, ax], fd))
        assert abs(gx[v, ax] - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(gx).max()), (v, ax, gx[v, ax], fd)


def test_batched_geometry_adjoint_equals_individual():
    """The batched sweep returns each case's own node-coordinate gradients."""
    from cfd2d import batch
    specs = [(5, 0.7, 1.0, 0.04), (8, 1.3, -2.0, 0.08), (6, 0.5, 4.0, 0.06)]
    cs = [cases.bump_channel(ny, height=h, mach=m, angle_deg=a, seed=ny, flux="LAX", gradient="GREEN_GAUSS",
                             limiter="NONE", rk_stages=2) for ny, m, a, h in specs]
    n = 10
    b = batch.BatchSolver(cs)
    mask = b.edge_mask("bump")
    res = b.adjoint_geometry(n, mask, ckpt_every=4)
    b.close()
    for c, (Jb, gUb, gbb, gxb) in zip(cs, res):
        s = native.GpuSolver(c)
        J, gU, gb, gx = s.adjoint_geometry(n, s.edge_mask("bump"), ckpt_every=4)
        s.close()
        assert abs(Jb - J) <= 1e-13 * abs(J)
        assert np.array_equal(gUb, gU)
        assert np.max(np.abs(gxb - gx)) <= 1e-13 * np.max(np.abs(gx))


def _host_vjp(mesh, lsq, Ge, Gc, lny):
    dxy = np.zeros(2 * mesh.nn)
    vp = lambda a: a.ctypes.data_as(native.C.c_void_p)
    native.check(native.lib().cfd2d_mesh_geometry_vjp(mesh._h, int(lsq), vp(Ge), vp(Gc), vp(lny), dxy))
    return dxy.reshape(-1, 2)


@pytest.mark.parametrize("kind,lsq", [("bump", False), ("bump", True), ("pulse", True), ("batch", False),
                                      ("batch", True)])
def test_geometry_vjp_device_equals_host(kind, lsq):
    """cfd2d_geometry_vjp_device (the node-coordinate reverse that
    cfd2d_adjoint_geometry runs on the GPU) gives the same bits as the host's
    cfd2d_mesh_geometry_vjp (checked against central differences of the mesh
    builder in test_geometry_adjoint.py) for random adjoints of every input:
    a bump channel, a periodic mesh (shifted neighbours) and a 3-case batch
    mesh; two calls on one context (the mesh's device copy reused), the
    second with lny left out."""
    from cfd2d import batch
    if kind == "bump":
        s = native.GpuSolver(bump("LAX", "LSQ" if lsq else "GREEN_GAUSS", 2, ny=9))
        mesh, ctx = s.mesh, s
    elif kind == "pulse":
        s = native.GpuSolver(pulse("LAX", "LSQ", 2, n=10))
        mesh, ctx = s.mesh, s
    else:
        cs = [cases.bump_channel(ny, height=0.05, seed=ny, flux="LAX", gradient="GREEN_GAUSS", limiter="NONE")
              for ny in (4, 7, 5)]
        s = batch.BatchSolver(cs)
        mesh, ctx = s.mesh, s.ctx
    rng = np.random.default_rng(11)
    ex = mesh.export()
    for call in range(2):
        Ge = rng.standard_normal((13, mesh.ne))
        Ge[8:12, ex["edge_int"][:, 3] < 0] = 0.0
        Gc = rng.standard_normal((7, mesh.nc))
        Gc[0, rng.random(mesh.nc) < 0.2] = 0.0  # cells without an area term
        lny = np.where(rng.random(mesh.ne) < 0.2, rng.standard_normal(mesh.ne), 0.0) if call == 0 else None
        host = _host_vjp(mesh, lsq, Ge, Gc, lny if lny is not None else np.zeros(mesh.ne))
        dev = ctx.geometry_vjp_device(lsq, Ge, Gc, lny)
        assert np.abs(host).max() > 1.0
        assert np.array_equal(dev, host), (call, np.abs(dev - host).max())
    s.close()


@pytest.mark.parametrize("flux,gradient,rk", [("LAX", "GREEN_GAUSS", 2), ("HLLC", "LSQ", 3)])
def test_adjoint_recording_is_the_same_for_every_fuse_mode(flux, gradient, rk):
    """The recorded gradient records come from the update pass (fuse modes 1
    and 2, the per-edge and tiled edge passes) or from the separate gradient
    pass (mode 0); all three give the same bits of J, dJ/dU0 and dJ/dpar."""
    c = bump(flux, gradient, rk, ny=10)
    res = []
    for mode in (1, 0, 2):
        g = native.GpuSolver(c)
        g.fuse_gradient(mode)
        res.append(g.adjoint(8, g.edge_mask("bump"), ckpt_every=3))
        g.close()
    for J, gU, gb in res[1:]:
        assert J == res[0][0]
        assert np.array_equal(gU, res[0][1])
        assert np.array_equal(gb, res[0][2])


def _draw_adjoint_case(rng, tube):
    flux = str(rng.choice(["LAX", "HLLC", "GODUNOV"]))
    gradient = str(rng.choice(["GREEN_GAUSS", "LSQ"]))
    rk = int(rng.choice([2, 3]))
    if not tube:
        c = cases.bump_channel(int(rng.integers(4, 8)), height=float(rng.uniform(0.02, 0.08)),
                               mach=float(rng.uniform(0.3, 2.0)), angle_deg=float(rng.uniform(-5, 5)),
                               seed=int(rng.integers(1, 1000)),
                               outlet=str(rng.choice(["BOUND_OUTLET", "BOUND_PRESSURE_OUTLET"])), flux=flux,
                               gradient=gradient, limiter="NONE", rk_stages=rk)
        kind = "bump"
    else:
        c = cases.shock_tube(int(rng.integers(8, 16)), int(rng.integers(3, 6)), seed=int(rng.integers(1, 1000)),
                             mild=True, flux=flux, gradient=gradient, limiter="NONE", rk_stages=rk)
        kind = "tube"
    return kind, c, int(rng.choice([0, 1, 3, 5]))


def test_seeded_adjoint_sweep_matches_tangent():
    """Seeded draws beyond the fixed list: LAX, HLLC or Godunov; Green-Gauss
    or LSQ; RK2 or RK3; bump channels with either outlet type and inflow
    Mach 0.3-2.0 at an angle, or (every third case) the mild shock tube (slip
    and no-slip walls, outlets); checkpoint intervals 0 (every stage
    recorded), 1, 3 or 5; the start state perturbed by 1e-3.  J within 1e-12
    and <dJ/dU0, v> + <dJ/dpar, w> within 1e-9 of the tangent model, as in
    the fixed list."""
    rng = np.random.default_rng(5)
    for it in range(15):
        kind, c, ck = _draw_adjoint_case(rng, tube=it % 3 == 2)
        cfg = c.config
        label = "case %d: %s %s %s RK%d ckpt %d cells %d" % (it, kind, cfg.flux, cfg.gradient, cfg.rk_stages, ck,
                                                          c.mesh.n_cells)
        print(label, flush=True)
        s = oracle.OracleSolver(c)
        tau, _ = s.tau()
        eg = s.geometry["edge_group"]
        m = c.mesh
        names = ["bump", "wall_bottom"] if kind == "bump" else ["edge_bottom"]  # J of order 1, no cancellation
        mask = np.isin(eg, [m.group_index(nm) for nm in names]).astype(np.int32)
        n = 10
        # a uniform start puts the two sides' wave speeds of every face at a tie
        # (LAX alpha, the HLLC bounds), where the map has no unique derivative and
        # the device's contracted reconstruction and the model can differ in the
        # last bit that breaks it: start from a seeded perturbation instead
        U0 = c.initial_state()
        U0 = U0 * (1.0 + 1e-3 * rng.standard_normal(U0.shape))
        U0[1:3] += 1e-3 * U0[0] * rng.standard_normal(U0[1:3].shape)
        g = native.GpuSolver(c)
        g.set_tau(tau)
        g.set_state(U0)
        J, gU, gb = g.adjoint(n, mask, ckpt_every=ck)
        g.close()
        v = rng.standard_normal(U0.shape) * np.abs(U0) * 1e-2
        tm = tangent.Model(s)
        w = rng.standard_normal(tm.bcpar0.shape) * np.abs(tm.bcpar0) * 1e-2
        Jm, dJ = tangent.tangent(s, U0, v, w, n, tau, mask)
        assert abs(J - Jm) <= 1e-12 * abs(Jm), label
        ad = float(np.sum(gU * v)) + float(np.sum(gb * w))
        assert abs(ad - dJ) <= 1e-9 * abs(dJ), (label, ad, dJ)
