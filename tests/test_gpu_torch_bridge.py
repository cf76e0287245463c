"""The torch.autograd wrapper of the adjoint (cfd2d/torch_adjoint.py) on the
GPU: autograd returns exactly what cfd2d_adjoint returns, and gradients in
the inflow parameters match central differences of the wrapped forward map
(h = 1e-6 relative, agreement 1e-6)."""
import numpy as np
import pytest

from cfd2d import case as cases
from cfd2d import native

pytestmark = pytest.mark.gpu


def _setup(ny=30):
    c = cases.bump_channel(ny, height=0.06, mach=0.8, seed=3, flux="HLLC", gradient="LSQ", limiter="NONE",
                           rk_stages=3)
    g = native.GpuSolver(c)
    par0 = np.array([list(b.params()) for b in c.config.boundaries if b.type != "BOUND_PERIODIC"])
    # the bottom wall meets the inlet and the outlet, so the boundary
    # parameters act on J within a few steps
    mask = g.edge_mask("bump") | g.edge_mask("wall_bottom")
    return c, g, mask, par0


def test_autograd_equals_c_abi():
    torch = pytest.importorskip("torch")
    from cfd2d import torch_adjoint
    c, g, mask, par0 = _setup()
    n = 15
    U0 = torch.tensor(c.initial_state(), dtype=torch.float64, requires_grad=True)
    par = torch.tensor(par0, dtype=torch.float64, requires_grad=True)
    J = torch_adjoint.wall_force_y(U0, par, g, n, mask, ckpt_every=4)
    gU, gp = torch.autograd.grad(J, (U0, par))
    g.set_bc_par(par0)
    g.set_state(c.initial_state())
    Jd, gUd, gpd = g.adjoint(n, mask, ckpt_every=4)
    assert float(J.detach()) == Jd
    assert np.array_equal(gU.numpy(), gUd)
    assert np.array_equal(gp.numpy(), gpd)
    g.close()


def test_inflow_gradient_matches_finite_differences():
    torch = pytest.importorskip("torch")
    from cfd2d import torch_adjoint
    c, g, mask, par0 = _setup()
    n = 20
    U0 = torch.tensor(c.initial_state(), dtype=torch.float64)
    par = torch.tensor(par0, dtype=torch.float64, requires_grad=True)
    J = torch_adjoint.wall_force_y(U0, par, g, n, mask)
    (gp,) = torch.autograd.grad(J, (par,))
    # inlet (u, v, T, p) and outlet pressure together
    w = np.zeros_like(par0)
    w[0] = [0.3, 0.1, 0.2, 0.4]
    w[1, 0] = 0.5
    w *= np.abs(par0)  # relative to each parameter
    h = 1e-6

    def Jat(t):
        with torch.no_grad():
            return float(torch_adjoint.wall_force_y(U0, torch.tensor(par0 + t * w), g, n, mask))

    fd = (Jat(h) - Jat(-h)) / (2 * h)
    J0 = Jat(0.0)
    ad = float(np.sum(gp.numpy() * w))
    g.close()
    print("inflow adjoint %.15e fd %.15e" % (ad, fd))
    assert abs(fd) > 1e-8 * abs(J0)  # the parameters do act on J
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)
