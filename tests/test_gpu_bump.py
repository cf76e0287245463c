"""The batched bump-channel workload (inlet from Mach and angle, pressure
outlet, slip walls and bump; exact Godunov flux, Green-Gauss gradients with
the minmod limiter, Heun RK2) on the GPU against the oracle, through the C
ABI: one case at a time, an uneven batch against each case's own oracle
run, and the full 256-case batch of the benchmark through properties that
hold at any size.

Tolerances: states 1e-10 relative L-inf per component after 40 steps (the
device differs from the CPU by FMA contraction and the last bit of its
roots), flags identical; the full batch: per case, the cells' residual sum
equals minus the case's boundary-edge fluxes to 1e-12 of the flux scale
(every interior flux cancels), and cases advanced in the batch are bitwise
the cases advanced alone.
"""
import numpy as np
import pytest

import oracle
from cfd2d import batch
from cfd2d import case as cases
from cfd2d import meshgen, native

pytestmark = pytest.mark.gpu


def _rel(a, b):
    den = np.max(np.abs(b), axis=1)
    den = np.where(den == 0, 1.0, den)
    return np.max(np.abs(a - b), axis=1) / den


@pytest.mark.parametrize("ny,mach,angle,height", [(10, 0.5, 0.0, 0.05), (9, 0.85, 2.0, 0.08), (8, 1.6, -3.0, 0.03),
                                                  (7, 2.7, 4.5, 0.1)], ids=str)
def test_bump_channel_matches_oracle(ny, mach, angle, height):
    c = cases.bump_channel(ny, height=height, mach=mach, angle_deg=angle, seed=ny)
    assert (c.config.flux, c.config.gradient, c.config.limiter, c.config.rk_stages) == \
        ("GODUNOV", "GREEN_GAUSS", "MINMOD", 2)
    s = native.GpuSolver(c)
    o = oracle.OracleSolver(c)
    assert abs(s.tau()[0] - o.tau()[0]) <= 1e-13 * o.tau()[0]
    R, Ro = s.residual(), o.residual()
    assert np.all(_rel(R, Ro) <= 1e-11), _rel(R, Ro)
    for n in (1, 9, 30):
        s.run(n)
        o.run(n)
        err = _rel(s.state(), o.state())
        assert np.all(err <= 1e-10), (n, err)
        np.testing.assert_array_equal(s.flags(), o.flags())
    assert s.status() == (0, 0) and o.newton_cap_hits() == 0
    # the flow has moved off the free stream everywhere near the bump
    assert np.abs(s.state() - c.initial_state()).max() > 1e-3
    s.close()


def test_uneven_batch_matches_per_case_oracle():
    specs = [(5, 0.3, -5.0, 0.02), (11, 0.8, 1.0, 0.1), (4, 1.2, 3.0, 0.06), (8, 2.0, -2.0, 0.04),
             (6, 2.8, 5.0, 0.09), (9, 0.5, 0.0, 0.07), (7, 1.6, -4.0, 0.03)]
    cs = [cases.bump_channel(ny, height=h, mach=m, angle_deg=a, seed=100 + k)
          for k, (ny, m, a, h) in enumerate(specs)]
    assert len({c.mesh.n_cells for c in cs}) == len(cs)
    b = batch.BatchSolver(cs)
    b.run(25)
    states = b.states()
    flags = b.ctx.flags()
    for k, c in enumerate(cs):
        o = oracle.OracleSolver(c)
        o.run(25)
        err = _rel(states[k], o.state())
        assert np.all(err <= 1e-10), (k, err)
        np.testing.assert_array_equal(flags[b.cell_off[k]:b.cell_off[k + 1]], o.flags())
        assert o.newton_cap_hits() == 0
    assert b.ctx.status() == (0, 0)
    b.close()


def test_full_batch_properties():
    """The 256-case batch (4.47 M cells, 2k-60k cells per case) after 20
    steps: finite, no Newton cap hit, per-case conservation of the residual,
    and three cases (smallest, median, largest) bitwise equal to themselves
    advanced alone."""
    specs = meshgen.bump_batch(256, seed=7)
    cs = [cases.bump_channel(s["ny"], height=s["height"], mach=s["mach"], angle_deg=s["angle_deg"],
                             seed=s["seed"]) for s in specs]
    b = batch.BatchSolver(cs)
    print("batch of %d cases, %d cells" % (len(cs), b.nc), flush=True)
    b.run(20)
    U = b.ctx.state()
    assert np.all(np.isfinite(U)) and b.ctx.status() == (0, 0)
    R = b.ctx.residual()
    F = b.ctx.edge_fluxes()
    ex = b.mesh.export()
    c2 = ex["edge_int"][:, 3]
    bnd = c2 < 0
    co, eo = b.cell_off, b.edge_off
    worst = 0.0
    for k in range(len(cs)):
        cell_sum = R[:, co[k]:co[k + 1]].sum(axis=1)
        e = slice(eo[k], eo[k + 1])
        bsum = F[e][bnd[e]].sum(axis=0)
        scale = np.abs(F[e]).sum(axis=0).max()
        worst = max(worst, float(np.abs(cell_sum + bsum).max() / scale))
    print("per-case residual balance, worst %.3e" % worst, flush=True)
    assert worst <= 1e-12
    sizes = np.array([c.mesh.n_cells for c in cs])
    order = np.argsort(sizes)
    for k in (order[0], order[len(order) // 2], order[-1]):
        s = native.GpuSolver(cs[k])
        s.run(20)
        assert np.array_equal(s.state(), U[:, co[k]:co[k + 1]]), k
        s.close()
    b.close()
