"""Derivatives with respect to the free-stream Mach number and flow angle.

A bump-channel case (cfd2d.case.bump_channel) derives everything from
(Mach, angle): the inlet parameters (u, v, T, p) = (M c cos a, M c sin a,
p/rho, p) with c = sqrt(gamma p / rho), rho = 1, p = 1/gamma, the pressure
outlet's p (independent of M and a), and the initial state (the free
stream in every cell).  Given the adjoint's dJ/dU0 and dJ/d(boundary
parameters), the chain rule through those relations gives dJ/dM and
dJ/da (per degree), with the time step held fixed as in the adjoint.
"""
from __future__ import annotations

import math

import numpy as np


def free_stream_derivatives(mach, angle_deg, gamma=1.4):
    """(u, v) and their derivatives d/dM and d/da (a in degrees)."""
    p, rho = 1.0 / gamma, 1.0
    c = math.sqrt(gamma * p / rho)
    a = math.radians(angle_deg)
    k = math.pi / 180.0
    u, v = mach * c * math.cos(a), mach * c * math.sin(a)
    return (u, v), (c * math.cos(a), c * math.sin(a)), (-mach * c * math.sin(a) * k, mach * c * math.cos(a) * k)


def mach_angle_gradient(case, dJ_dU0, dJ_dpar, inlet="inlet"):
    """(dJ/dMach, dJ/dangle_deg) of a bump_channel case from the adjoint's
    dJ/dU0 (4, nc) and dJ/dpar (n_bc, 4)."""
    mach, ang = case.meta["mach"], case.meta["angle_deg"]
    gamma = case.config.materials[0].gamma
    (u, v), (du_m, dv_m), (du_a, dv_a) = free_stream_derivatives(mach, ang, gamma)
    bcs = [b for b in case.config.boundaries if b.type != "BOUND_PERIODIC"]
    k = [b.name for b in bcs].index(inlet)
    gp = np.asarray(dJ_dpar)
    U0 = case.initial_state()
    rho = U0[0]
    g = np.asarray(dJ_dU0)

    def through(du, dv):
        # inlet (u, v) and the free-stream initial state (rho u, rho v, E)
        s = gp[k, 0] * du + gp[k, 1] * dv
        s += float(np.sum(g[1] * rho * du) + np.sum(g[2] * rho * dv) + np.sum(g[3] * rho * (u * du + v * dv)))
        return s

    return through(du_m, dv_m), through(du_a, dv_a)
