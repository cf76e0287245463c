"""HIP path vs the reference (golden vectors) and vs the oracle, through the C ABI.

Tolerances (FP64 on both sides; the device contracts a*b+c into FMA and its
exp/log differ from glibc's in the last bit, so agreement is to round-off,
not bitwise):
  time step, gradients, edge fluxes      1e-12 relative to the component's max
  stage-1 residual                       |dR| / max(|R|, sum_f |F_f|_inf l/2) <= 1e-12
  state after 1..100 steps               1e-11 relative L-inf per component
  state after 1000 steps                 1e-9  relative L-inf per component
"""
import numpy as np
import pytest

import golden_cases as gc
import oracle
from cfd2d import case as cases
from cfd2d import native

pytestmark = pytest.mark.gpu


def _rel(a, b, axis=None):
    den = np.max(np.abs(b), axis=axis)
    den = np.where(den == 0, 1.0, den)
    return np.max(np.abs(a - b), axis=axis) / den


@pytest.fixture(scope="module", params=gc.CASES)
def golden(request):
    name = request.param
    return name, gc.case(name), gc.load(name)


def test_device_present():
    assert native.device_count() >= 1


def test_time_step(golden):
    name, case, g = golden
    s = native.GpuSolver(case)
    tau, ct = s.tau()
    assert abs(tau - g["tau"][0]) <= 1e-13 * g["tau"][0]
    assert np.max(np.abs(ct - g["tau"][1:]) / g["tau"][1:]) <= 1e-13
    s.close()


def test_gradients(golden):
    name, case, g = golden
    s = native.GpuSolver(case)
    G = s.gradients()
    assert np.all(_rel(G, g["grad0"], axis=1) <= 1e-12), _rel(G, g["grad0"], axis=1)
    s.close()


def test_edge_fluxes(golden):
    name, case, g = golden
    s = native.GpuSolver(case)
    F = s.edge_fluxes()
    assert np.all(_rel(F, g["edgeflux0"], axis=0) <= 1e-12), _rel(F, g["edgeflux0"], axis=0)
    s.close()


def residual_scale(cell_edges, edgeflux):
    """Per cell: sum over its faces of the face-flux vector's max-norm |F_f|_inf l_f/2.

    The round-off of every residual component is set by the largest flux
    components (pressure terms), so the scale of a component whose fluxes
    vanish in exact arithmetic (mass flux of a fluid at rest) is that of
    the whole flux vector, not of its own noise."""
    fmax = np.abs(edgeflux).max(axis=1)
    return fmax[cell_edges].sum(axis=1)


def test_residual_1e12(golden):
    """The one-step residual gate: per-cell residual within 1e-12 relative,
    |R_gpu - R_ref| / max(|R_ref|, sum_f |F_f|_inf l_f/2) <= 1e-12."""
    name, case, g = golden
    s = native.GpuSolver(case)
    R = s.residual()
    ref = g["residual0"]
    scale = residual_scale(s.mesh.export()["cell_edges"], g["edgeflux0"])
    den = np.maximum(np.abs(ref), scale[None, :])
    err = np.max(np.abs(R - ref) / den)
    assert err <= 1e-12, err
    s.close()


def test_steps(golden):
    name, case, g = golden
    s = native.GpuSolver(case)
    steps = sorted(int(k.split("step")[1]) for k in g.files if k.startswith("u_step"))
    done = 0
    for st in steps:
        s.run(st - done)
        done = st
        U = s.state()
        tol = 1e-9 if st >= 1000 else 1e-11
        err = _rel(U, g["u_step%d" % st], axis=1)
        assert np.all(err <= tol), (name, st, err)
        np.testing.assert_array_equal(s.flags(), g["flags_step%d" % st])
    assert s.status() == (0, 0)
    s.close()


def test_deterministic_and_graph_replay():
    """Two eager runs and one hipGraph run agree bit-for-bit."""
    case = gc.case("inlet")
    outs = []
    for graph in (False, False, True):
        s = native.GpuSolver(case)
        s.run(7, graph=graph)
        s.run(5, graph=graph)
        outs.append(s.state())
        s.close()
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0], outs[2])


@pytest.mark.parametrize("opts", [dict(), dict(gradient="LSQ", rk_stages=3, flux="HLLC"),
                                  dict(viscous=True, gradient="LSQ", rk_stages=3)], ids=str)
def test_fused_gradient_edge_pass_is_bitwise(opts):
    """The edge pass that forms the cell gradients itself gives the same bits
    as the separate gradient pass (inlet, walls, outlet; inviscid and NS)."""
    case = gc.case("inlet", **opts)
    outs = []
    for fuse in (True, False):
        s = native.GpuSolver(case)
        s.fuse_gradient(fuse)
        s.run(25)
        outs.append(s.state())
        s.close()
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("name,opts", [("inlet", dict()), ("bounds", dict(limiter="MINMOD")),
                                       ("twomat", dict(gradient="LSQ", rk_stages=3))], ids=str)
def test_streamed_state_is_bitwise(name, opts):
    """Non-temporal loads and stores of the state in the update pass change
    only where the data is cached, not the values (RK2 and RK3 stages, LIM
    flags, two materials)."""
    case = gc.case(name, **opts)
    outs = []
    for mode in (True, False, None):
        s = native.GpuSolver(case)
        s.stream_state(mode)
        s.run(25)
        outs.append((s.state(), s.flags()))
        s.close()
    for o in outs[1:]:
        assert np.array_equal(outs[0][0], o[0]) and np.array_equal(outs[0][1], o[1])


@pytest.mark.parametrize("name,opts", [("inlet", dict()), ("inlet", dict(gradient="LSQ", rk_stages=3, flux="HLLC")),
                                       ("inlet", dict(viscous=True, gradient="LSQ", rk_stages=3)),
                                       ("bounds", dict(limiter="MINMOD")), ("twomat", dict(gradient="LSQ", rk_stages=3)),
                                       ("tiny", dict(flux="LAX"))], ids=str)
def test_tiled_edge_pass_is_bitwise(name, opts):
    """The tiled edge pass (gradients formed once per tile and staged in LDS)
    gives the same states, flags and edge fluxes as the per-edge pass
    (inlet, outlet, walls, two materials, the limiter, four fluxes)."""
    case = gc.case(name, **opts)
    outs = []
    for mode in (1, 2):
        s = native.GpuSolver(case)
        s.fuse_gradient(mode)
        s.run(25)
        outs.append((s.state(), s.flags(), s.edge_fluxes()))
        s.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("order", ["generator", "shuffle"])
def test_tiled_edge_pass_periodic_and_unordered(order):
    """Periodic Navier-Stokes (the benchmark's scheme at 8 192 cells), in the
    generator's cell order and shuffled (whose tiles stage more cells, so
    the library picks smaller ones), and a batch of bump channels: the tiled
    pass equals the per-edge pass bitwise."""
    from cfd2d import batch
    c = cases.periodic_pulse(64, seed=5, gradient="LSQ", rk_stages=3, viscous=True, cfl_mode="SPECTRAL")
    if order == "shuffle":
        from cfd2d import renumber
        c, _ = renumber.renumber_case(c, "random", seed=3)
    outs = []
    for mode in (1, 2):
        s = native.GpuSolver(c)
        s.fuse_gradient(mode)
        s.run(10)
        outs.append(s.state())
        s.close()
    assert np.array_equal(outs[0], outs[1])
    cs = [cases.bump_channel(ny, height=0.05, mach=m, seed=ny, limiter="MINMOD") for ny, m in ((5, 0.7), (9, 1.6), (7, 2.4))]
    outs = []
    for mode in (1, 2):
        b = batch.BatchSolver(cs)
        b.ctx.fuse_gradient(mode)
        b.run(10)
        outs.append(b.ctx.state())
        b.close()
    assert np.array_equal(outs[0], outs[1])


def test_fused_limited_gradient_is_bitwise():
    """With the minmod limiter too (mode 1 forms and limits the gradients in
    the edge pass; the default runs the separate pass)."""
    case = cases.shock_tube(40, 6, seed=4, mild=False, limiter="MINMOD", max_newton=200)
    outs = []
    for mode in (True, None):
        s = native.GpuSolver(case)
        s.fuse_gradient(mode)
        s.run(30)
        outs.append((s.state(), s.flags()))
        s.close()
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("name,opts", [("inlet", dict()), ("bounds", dict(limiter="MINMOD")), ("steady", dict()),
                                       ("twomat", dict(gradient="LSQ", rk_stages=3))], ids=str)
def test_autotune_restores_state_and_keeps_results(name, opts):
    case = gc.case(name, **opts)
    a = native.GpuSolver(case)
    b = native.GpuSolver(case)
    a.run(7)
    b.run(7)
    U, F = a.state(), a.flags()
    choice = a.autotune(trial_steps=2)
    assert choice["fuse"] in (0, 1, 2) and choice["graph"] in (0, 1) and choice["stream"] in (0, 1)
    assert choice["edge_shape"] in range(4)
    assert choice["ms_per_step"] > 0
    assert np.array_equal(a.state(), U) and np.array_equal(a.flags(), F)
    assert a.status() == b.status()
    a.run(20)
    b.run(20)
    assert np.array_equal(a.state(), b.state()) and np.array_equal(a.flags(), b.flags())
    a.close()
    b.close()


EXT = [  # (full Sod state?, options): unlimited variants on the mild tube, limited ones on Sod
    (False, dict(flux="LAX")), (False, dict(flux="HLLC")), (False, dict(flux="ROE")),
    (False, dict(gradient="LSQ")), (False, dict(rk_stages=3)), (False, dict(cfl_mode="SPECTRAL", cfl=0.4)),
    (False, dict(gradient="LSQ", rk_stages=3, flux="HLLC")), (False, dict(steady=True, cfl=0.4)),
    (True, dict(limiter="MINMOD")), (True, dict(limiter="MINMOD", flux="ROE")),
    (True, dict(limiter="MINMOD", gradient="LSQ", rk_stages=3, flux="HLLC", cfl_mode="SPECTRAL", cfl=0.4)),
]


@pytest.mark.parametrize("sod,opts", EXT, ids=lambda o: str(o))
def test_extensions_match_oracle(sod, opts):
    case = cases.shock_tube(40, 6, seed=4, mild=not sod, max_newton=200, **opts)
    s = native.GpuSolver(case)
    o = oracle.OracleSolver(case)
    assert abs(s.tau()[0] - o.tau()[0]) <= 1e-13 * o.tau()[0]
    R = s.residual()
    Ro = o.residual()
    assert np.all(_rel(R, Ro, axis=1) <= 1e-11), _rel(R, Ro, axis=1)
    s.run(20)
    o.run(20)
    assert np.all(_rel(s.state(), o.state(), axis=1) <= 1e-10), _rel(s.state(), o.state(), axis=1)
    assert np.array_equal(s.flags(), o.flags())
    assert s.status() == (0, 0) and o.newton_cap_hits() == 0
    s.close()


def test_viscous_periodic_matches_oracle():
    case = cases.periodic_pulse(24, seed=5)
    s = native.GpuSolver(case)
    o = oracle.OracleSolver(case)
    assert abs(s.tau()[0] - o.tau()[0]) <= 1e-13 * o.tau()[0]
    assert np.all(_rel(s.residual(), o.residual(), axis=1) <= 1e-11)
    s.run(10)
    o.run(10)
    assert np.all(_rel(s.state(), o.state(), axis=1) <= 1e-11)
    s.close()


def test_periodic_ns_1000_steps_matches_oracle():
    """The benchmark's scheme over the benchmark's horizon: periodic
    Navier-Stokes (LSQ gradients, exact Godunov flux, SSP-RK3, spectral dt,
    mu = 1e-4, Pr = 0.72) on 18 432 cells for 1 000 steps against the
    oracle: conserved variables within 1e-9 relative L-inf per component
    (the benchmark's parity target), checked after 10, 100 and 1 000 steps."""
    case = cases.periodic_pulse(96, seed=2)
    s = native.GpuSolver(case)
    o = oracle.OracleSolver(case)
    done = 0
    for n, tol in ((10, 1e-11), (100, 1e-11), (1000, 1e-9)):
        s.run(n - done)
        o.run(n - done)
        done = n
        err = _rel(s.state(), o.state(), axis=1)
        assert np.all(err <= tol), (n, err)
    print("periodic NS 18 432 cells after 1000 steps: max rel err per component", err)
    assert s.status() == (0, 0) and o.newton_cap_hits() == 0
    s.close()


def test_periodic_conservation():
    """Periodic torus: sum S*U is conserved to round-off."""
    case = cases.periodic_pulse(32, seed=6)
    s = native.GpuSolver(case)
    S = s.mesh.export()["cell_geom"][:, 0]
    tot0 = (s.state() * S).sum(axis=1)
    s.run(100)
    tot1 = (s.state() * S).sum(axis=1)
    assert np.all(np.abs(tot1 - tot0) <= 1e-13 * np.abs(tot0).max())
    s.close()


def test_benchmark_size_conservation_and_variants():
    """The benchmark's own case (2,097,152-cell periodic Navier-Stokes, LSQ,
    SSP-RK3, state streamed by default at this size): over 20 steps sum S*U
    is conserved to round-off, nothing is capped or non-finite, and the
    variants with the gradients in their own pass and the state cached give
    the same bits."""
    case = cases.periodic_pulse(1024, seed=2)
    s = native.GpuSolver(case)
    S = s.mesh.export()["cell_geom"][:, 0]
    U0 = s.state()
    tot0 = (U0 * S).sum(axis=1)
    s.run(20)
    U1 = s.state()
    assert np.all(np.isfinite(U1)) and s.status() == (0, 0)
    tot1 = (U1 * S).sum(axis=1)
    assert np.all(np.abs(tot1 - tot0) <= 1e-12 * np.abs(tot0).max())
    s.set_state(U0)
    s.fuse_gradient(False)
    s.stream_state(False)
    s.run(20)
    assert np.array_equal(s.state(), U1)
    s.close()


def test_benchmark_size_matches_oracle():
    """The benchmark's own case at full size (2,097,152 cells, periodic
    Navier-Stokes, LSQ, exact Godunov, SSP-RK3, spectral dt) against the
    oracle: the residual of the initial state within the 1e-12 gate of
    test_residual_1e12 and the states after two steps within
    1e-11 relative per component; the tiled edge pass gives the same bits."""
    case = cases.periodic_pulse(1024, seed=2)
    s = native.GpuSolver(case)
    o = oracle.OracleSolver(case)
    assert abs(s.tau()[0] - o.tau()[0]) <= 1e-13 * o.tau()[0]
    R, Ro = s.residual(), o.residual()
    den = np.maximum(np.abs(Ro), residual_scale(s.mesh.export()["cell_edges"], o.edge_fluxes())[None, :])
    rerr = float(np.max(np.abs(R - Ro) / den))
    assert rerr <= 1e-12, rerr
    s.run(2)
    o.run(2)
    err = _rel(s.state(), o.state(), axis=1)
    print("2.1 M cells: residual %.2e, after 2 steps max rel err per component %s" % (rerr, err))
    assert np.all(err <= 1e-11), err
    U2 = s.state()
    s.set_state(case.initial_state())
    s.fuse_gradient(2)
    s.run(2)
    assert np.array_equal(s.state(), U2)
    s.close()


@pytest.mark.parametrize("nx,ny", [(1, 1), (2, 1), (3, 2)])
@pytest.mark.parametrize("opts", [dict(), dict(limiter="MINMOD", gradient="LSQ", rk_stages=3),
                                  dict(viscous=True, flux="HLLC", cfl_mode="SPECTRAL", cfl=0.4)], ids=str)
def test_tiny_meshes_match_oracle(nx, ny, opts):
    """Meshes of 2 to 12 cells, where every cell touches the boundary (the
    inlet, the outlet and both walls meet in one cell, or every edge of a
    cell is a boundary edge) and every launch is a fraction of one
    workgroup: 15 steps against the oracle for the three gradient
    placements (separate pass, per-edge, tiled), which agree bitwise."""
    import dataclasses
    from cfd2d import meshgen
    base = gc.case("inlet", **opts)
    case = dataclasses.replace(base, mesh=meshgen.rect_mesh(nx=nx, ny=ny, lx=1.0, ly=0.2, seed=3))
    o = oracle.OracleSolver(case)
    o.run(15)
    outs = []
    for mode in (0, 1, 2):
        s = native.GpuSolver(case)
        s.fuse_gradient(mode)
        s.run(15)
        outs.append(s.state())
        assert s.status() == (0, 0)
        np.testing.assert_array_equal(s.flags(), o.flags())
        s.close()
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    err = _rel(outs[0], o.state(), axis=1)
    assert np.all(err <= 1e-11), err


def test_dense_remediation_matches_oracle():
    """Many LIM cells at once: the bounds case with p_max = 0.99 on an
    8 000-cell tube flags the whole high-pressure half (~4 000 cells, up to
    201 of a 256-cell block, next to each other: a cell then averages
    neighbours remediated earlier in the same sweep), through the 32-step
    release.  States and flags equal the oracle's sequential sweep after
    every step."""
    import dataclasses
    from cfd2d import meshgen
    case = dataclasses.replace(gc.case("bounds", p_max=0.99),
                               mesh=meshgen.rect_mesh(nx=200, ny=20, lx=1.0, ly=0.2, seed=9))
    s = native.GpuSolver(case)
    o = oracle.OracleSolver(case)
    most = 0
    for st in range(40):
        s.run(1)
        o.run(1)
        np.testing.assert_array_equal(s.flags(), o.flags())
        assert np.all(_rel(s.state(), o.state(), axis=1) <= 1e-11), st
        most = max(most, int(np.count_nonzero(s.flags() & 0x2)))
    print("most LIM cells in one step: %d" % most)
    assert most >= 3000
    s.close()


def test_sod_tube_config_matches_oracle():
    """The CPU-runnable benchmark configuration end to end: the 8 000-cell
    jittered tube with the full Sod state, exact Godunov flux, Green-Gauss
    gradients with the minmod limiter, Heun RK2, CFL 0.5, 1 000 steps, on the
    GPU against the oracle.  States within 1e-9 relative L-inf per component
    (the tolerance of the unlimited 1 000-step golden gate), flags equal,
    no Newton cap hit."""
    case = cases.shock_tube(200, 20, seed=1, mild=False, limiter="MINMOD", cfl=0.5, max_newton=500)
    s = native.GpuSolver(case)
    o = oracle.OracleSolver(case)
    for n in (10, 90, 900):
        s.run(n)
        o.run(n)
        err = _rel(s.state(), o.state(), axis=1)
        assert np.all(err <= 1e-9), (n, err)
        np.testing.assert_array_equal(s.flags(), o.flags())
    assert s.status() == (0, 0) and o.newton_cap_hits() == 0
    print("Sod tube after 1000 steps: max rel err per component", _rel(s.state(), o.state(), axis=1))
    s.close()


def test_sod_full_with_minmod_runs_and_bounds():
    """The full Sod state (which hangs the reference's unlimited path) runs
    with the minmod limiter; density stays within the initial bounds."""
    case = cases.shock_tube(200, 20, seed=1, mild=False, limiter="MINMOD")
    s = native.GpuSolver(case)
    s.run(200)
    U = s.state()
    assert U[0].min() >= 0.125 - 1e-12 and U[0].max() <= 1.0 + 1e-12
    assert s.status() == (0, 0)
    s.close()


def test_newton_cap_reports_error():
    """Unlimited Sod makes the reconstructed pressure negative: the bounded
    Newton loop stops at its cap and the context reports CFD2D_ERR_NEWTON."""
    case = cases.shock_tube(10, 2, seed=5, mild=False, max_newton=50)
    s = native.GpuSolver(case)
    s.step(1)
    with pytest.raises(native.Cfd2dError) as e:
        s.sync()
    assert e.value.code == 3
    s.close()


def test_batched_equals_individual():
    """Concatenated cases advance exactly as when run one by one."""
    cs = [cases.bump_channel(ny, height=h, mach=m, seed=k) for k, (ny, h, m) in
          enumerate([(6, 0.05, 0.5), (8, 0.08, 1.5), (5, 0.03, 2.5)])]
    singles = []
    for c in cs:
        s = native.GpuSolver(c)
        s.run(15)
        singles.append(s.state())
        s.close()
    from cfd2d import batch
    b = batch.BatchSolver(cs)
    b.run(15)
    states = b.states()
    for a, ref in zip(states, singles):
        assert np.array_equal(a, ref)
    b.close()


@pytest.mark.parametrize("opts", [dict(), dict(gradient="LSQ", rk_stages=3, viscous=True)], ids=str)
def test_edge_launch_shapes_are_bitwise(opts):
    """Every launch shape of the edge pass (threads per workgroup, waves per
    SIMD) gives the same bits, fused or stored gradients."""
    case = gc.case("inlet", **opts)
    outs = []
    for shape in range(4):
        for fuse in (True, False):
            s = native.GpuSolver(case)
            s.edge_shape(shape)
            s.fuse_gradient(fuse)
            s.run(15)
            outs.append(s.state())
            assert s.tuning()["edge_shape"] == shape
            s.close()
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)
