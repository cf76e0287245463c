"""Runs the adjoint cases of the host sanitizer check with the library named
by CFD2D_LIB (a host build of libcfd2d, loaded under AddressSanitizer) and
writes every output to the .npz file named on the command line.

Cases: bump channels (inlet, pressure outlet, slip walls) and a periodic
pulse, over the LAX, HLLC and exact Godunov fluxes, Green-Gauss and
least-squares gradients, Heun RK2 and SSP-RK3; for each, cfd2d_adjoint and
cfd2d_adjoint_geometry (the geometry-sensitivity kernels) with a checkpoint
interval that makes the sweep recompute segments.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "cfd-2d_amd"), os.path.join(REPO, "oracle")]

from cfd2d import case as cases  # noqa: E402
from cfd2d import native  # noqa: E402

NSTEPS = 8

CASES = [("bump", "LAX", "GREEN_GAUSS", 2), ("bump", "HLLC", "LSQ", 3), ("bump", "LAX", "LSQ", 2),
         ("bump", "GODUNOV", "GREEN_GAUSS", 2), ("pulse", "HLLC", "LSQ", 2)]


def make(kind, flux, gradient, rk):
    if kind == "bump":
        return cases.bump_channel(6, height=0.06, mach=0.8, seed=3, flux=flux, gradient=gradient, limiter="NONE",
                                  rk_stages=rk)
    return cases.periodic_pulse(12, seed=2, flux=flux, gradient=gradient, rk_stages=rk, viscous=False,
                                cfl_mode="REFERENCE")


def mask_of(g, kind):
    if kind == "bump":
        return (g.edge_mask("bump") | g.edge_mask("wall_bottom")).astype(np.int32)
    c1 = g.mesh.export()["edge_int"][:, 2]
    return (c1 < g.nc // 4).astype(np.int32)


def main(out):
    res = {}
    for kind, flux, gradient, rk in CASES:
        name = "%s_%s_%s_%d" % (kind, flux, gradient, rk)
        print(name, flush=True)
        c = make(kind, flux, gradient, rk)
        g = native.GpuSolver(c)
        tau, _ = g.tau()
        mask = mask_of(g, kind)
        J, gU, gb = g.adjoint(NSTEPS, mask, ckpt_every=3)
        res[name + "/J"], res[name + "/gU"], res[name + "/gb"] = np.array(J), gU, gb
        if kind == "bump":
            g.set_state(c.initial_state())
            J, gU, gb, gx = g.adjoint_geometry(NSTEPS, mask, ckpt_every=3)
            res[name + "/geo_J"], res[name + "/geo_gU"] = np.array(J), gU
            res[name + "/geo_gb"], res[name + "/geo_gx"] = gb, gx
        res[name + "/tau"] = np.array(tau)
        g.close()
    np.savez(out, **res)


if __name__ == "__main__":
    main(sys.argv[1])
