"""The implicit method's matrix fill on the device (cfd2d_implicit_assemble)
against the oracle's restatement of FVM_TVD_IMPLICIT::run
(src/methods/fvm_tvd_implicit.cpp:534-663), through the C ABI.

Tolerances (FP64 both sides; the device contracts a*b+c into FMA):
  block columns                  exact
  Jacobian blocks                1e-12 relative to the largest entry of the cell's row
  right-hand side                1e-12 relative to the largest |rhs| of the component
  ctau                           1e-13 relative
At sizes the oracle does not reach, size-independent properties: on a
periodic mesh the block columns sum to zero once S/tau is removed, and the
right-hand side equals the explicit residual (to 1e-12 of the per-cell sum
of face-flux magnitudes: the two edge kernels round differently).
"""
import numpy as np
import pytest

import golden_cases as gc
import oracle
from cfd2d import case as cases
from cfd2d import native

pytestmark = pytest.mark.gpu


def compare(s, o, steady, cfl=0.6, tau=2.5e-3):
    B, C, R, T = s.implicit_assemble(steady=steady, cfl=cfl, tau=tau)
    Bo, Co, Ro, To = o.implicit_assemble(steady=steady, cfl=cfl, tau=tau)
    np.testing.assert_array_equal(C, Co)
    rowmax = np.abs(Bo).reshape(len(Bo), -1).max(axis=1)
    err = np.abs(B - Bo).reshape(len(Bo), -1).max(axis=1)
    assert np.all(err <= 1e-12 * rowmax), (err / rowmax).max()
    den = np.abs(Ro).max(axis=0)
    assert np.all(np.abs(R - Ro).max(axis=0) <= 1e-12 * den), np.abs(R - Ro).max(axis=0) / den
    np.testing.assert_allclose(T, To, rtol=1e-13)
    return B, C, R, T


@pytest.mark.parametrize("flux", ["GODUNOV", "LAX"])
@pytest.mark.parametrize("name", gc.CASES)
def test_golden_cases_match_oracle(name, flux):
    case = gc.case(name, flux=flux)
    s = native.GpuSolver(case)
    o = oracle.OracleSolver(case)
    steady = bool(gc.spec(name)["steady"])
    compare(s, o, steady)
    # after some explicit steps (a non-trivial state everywhere)
    s.run(10)
    o.set_state(s.state())
    compare(s, o, steady)
    compare(s, o, not steady)
    assert s.status() == (0, 0)
    s.close()


@pytest.mark.parametrize("mach,flux", [(0.6, "GODUNOV"), (2.0, "GODUNOV"), (0.6, "LAX"), (2.0, "LAX")])
def test_bump_channel_matches_oracle(mach, flux):
    case = cases.bump_channel(10, height=0.08, mach=mach, angle_deg=3.0, outlet="BOUND_OUTLET", flux=flux,
                              limiter="NONE", max_newton=200)
    s = native.GpuSolver(case)
    o = oracle.OracleSolver(case)
    s.run(25)
    o.set_state(s.state())
    for steady in (True, False):
        compare(s, o, steady)
    s.close()


def test_periodic_matches_oracle_and_explicit_residual():
    case = cases.periodic_pulse(16, seed=3, flux="GODUNOV", gradient="GREEN_GAUSS", viscous=False, rk_stages=2,
                                cfl_mode="REFERENCE")
    s = native.GpuSolver(case)
    o = oracle.OracleSolver(case)
    _, _, R, _ = compare(s, o, steady=True)
    Rexp = s.residual()
    assert np.all(np.abs(R.T - Rexp) <= 1e-12 * flux_scale(s))
    s.close()


def flux_scale(s):
    """Per cell: sum over its faces of |F_f|_inf l_f/2 (the explicit edge fluxes),
    the size of the terms whose cancellation forms the residual."""
    F = s.edge_fluxes()
    ce = s.mesh.export()["cell_edges"]
    return np.abs(F).max(axis=1)[ce].sum(axis=1)


def column_sums(B, C, S, ctau):
    nc = len(B)
    Bn = B.copy()
    Bn[:, 0] -= (S / ctau)[:, None, None] * np.eye(4)
    out = np.zeros((nc, 4, 4))
    for k in range(4):
        ok = C[:, k] >= 0
        np.add.at(out, C[ok, k], Bn[ok, k])
    return out, np.abs(Bn).max()


def test_large_periodic_properties():
    """131k cells: column sums vanish, RHS = explicit residual, steady ctau
    positive and finite."""
    case = cases.periodic_pulse(256, seed=4, flux="GODUNOV", gradient="GREEN_GAUSS", viscous=False, rk_stages=2,
                                cfl_mode="REFERENCE")
    s = native.GpuSolver(case)
    B, C, R, T = s.implicit_assemble(steady=True, cfl=0.8)
    S = s.mesh.export()["cell_geom"][:, 0]
    cs, m = column_sums(B, C, S, T)
    assert np.abs(cs).max() <= 1e-13 * m
    Rexp = s.residual()
    assert np.all(np.abs(R.T - Rexp) <= 1e-12 * flux_scale(s))
    assert np.all(np.isfinite(T)) and np.all(T > 0)
    s.close()


def test_assembly_leaves_explicit_stepping_unchanged():
    case = gc.case("inlet")
    a = native.GpuSolver(case)
    b = native.GpuSolver(case)
    a.run(6)
    b.run(3)
    b.implicit_assemble(steady=False, tau=1e-3)
    b.run(3)
    assert np.array_equal(a.state(), b.state())
    a.close()
    b.close()


def test_rejects_pressure_outlet_and_other_fluxes():
    s = native.GpuSolver(cases.bump_channel(4, limiter="NONE"))
    with pytest.raises(native.Cfd2dError):
        s.implicit_assemble()
    s.close()
    s = native.GpuSolver(gc.case("tiny", flux="HLLC"))
    with pytest.raises(native.Cfd2dError):
        s.implicit_assemble()
    with pytest.raises(native.Cfd2dError):
        s.implicit_assemble(steady=False, tau=0.0)
    s.close()


@pytest.mark.parametrize("flux,steady", [("GODUNOV", True), ("LAX", True), ("GODUNOV", False)])
def test_implicit_run_with_device_assembly_matches_oracle_assembly(flux, steady):
    """The implicit run loop (host GMRES) fed by the device's matrix fill
    follows the one fed by the oracle's, step for step."""
    from cfd2d import implicit

    c = cases.bump_channel(10, height=0.06, mach=0.6, outlet="BOUND_OUTLET", flux=flux, limiter="NONE",
                           max_newton=200)
    o = oracle.OracleSolver(c)

    def oracle_fill(U, st, cfl, tau):
        o.set_state(U)
        return o.implicit_assemble(steady=st, cfl=cfl, tau=tau)

    kw = dict(steady=steady, cfl=1.0, tau=5e-3)
    g = implicit.ImplicitSolver(c, **kw)
    h = implicit.ImplicitSolver(c, assemble=oracle_fill, **kw)
    for _ in range(12):
        assert g.step() == h.step() == implicit.RESULT_OK
        assert g.last_iters == h.last_iters
        den = np.abs(h.U).max(axis=1)
        assert np.all(np.abs(g.U - h.U).max(axis=1) <= 1e-10 * den)
    g.close()


def test_seeded_assembly_sweep_matches_oracle():
    """Seeded draws beyond the fixed cases: the mild tube or a bump channel
    (outlet, inflow Mach 0.3-2.5 at an angle), Godunov or LAX, steady or
    unsteady, CFL 0.3-1.0 and tau 1e-3-5e-3, each from a seeded 1e-2
    perturbation of its state (set on both sides); the fill within the
    tolerances above."""
    rng = np.random.default_rng(9)
    for it in range(20):
        flux = str(rng.choice(["GODUNOV", "LAX"]))
        if it % 2:
            case = cases.shock_tube(int(rng.integers(8, 20)), int(rng.integers(3, 6)), seed=int(rng.integers(1, 999)),
                                    mild=True, flux=flux, limiter="NONE", max_newton=200)
        else:
            case = cases.bump_channel(int(rng.integers(5, 12)), height=float(rng.uniform(0.02, 0.1)),
                                      mach=float(rng.uniform(0.3, 2.5)), angle_deg=float(rng.uniform(-5, 5)),
                                      seed=int(rng.integers(1, 999)), outlet="BOUND_OUTLET", flux=flux,
                                      limiter="NONE", max_newton=200)
        steady = bool(rng.integers(0, 2))
        cfl, tau = float(rng.uniform(0.3, 1.0)), float(rng.uniform(1e-3, 5e-3))
        print("case %d: %s cells %d steady %s cfl %.3f tau %.2e" % (it, flux, case.mesh.n_cells, steady, cfl, tau),
              flush=True)
        s = native.GpuSolver(case)
        o = oracle.OracleSolver(case)
        U = s.state()
        U = U * (1.0 + 1e-2 * rng.standard_normal(U.shape))
        s.set_state(U)
        o.set_state(U)
        compare(s, o, steady, cfl=cfl, tau=tau)
        s.close()
