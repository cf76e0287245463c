"""A case = mesh + configuration + initial state, resolved to index arrays.

Name resolution follows FVM_TVD::init: a boundary edge gets the first
<boundCond> whose name equals its group name (src/methods/fvm_tvd.cpp:132-174,
an unknown name is fatal), and a cell gets the region whose name equals its
cell group (getRegionByName, fvm_tvd.cpp:773-781), whose material index it
uses (fvm_tvd.cpp:788-792) and whose state it starts from (199-204).
"""
from __future__ import annotations

from dataclasses import dataclass, field
import math

import numpy as np

from . import meshgen
from .config import Boundary, Material, Region, SolverConfig, prim_to_cons, region_state


@dataclass
class Case:
    mesh: meshgen.TriMesh
    config: SolverConfig
    U0: np.ndarray = None          # (4, N) conserved initial state, SoA
    name: str = "case"
    meta: dict = field(default_factory=dict)

    def group_bc(self):
        """bc index per edge group (-1 for periodic or unmapped groups)."""
        return np.array([self.config.bc_index(g) for g in self.mesh.edge_group_names], np.int32)

    def cell_material(self):
        regs = {r.name: r for r in self.config.regions}
        mats = []
        for g in self.mesh.cell_group_names:
            if g in regs:
                mats.append(regs[g].material)
            else:
                mats.append(0)
        return np.asarray(mats, np.int32)[self.mesh.cell_group]

    def periodic_pairs(self):
        """(group_a, group_b, sx, sy) index tuples of the periodic extension."""
        out = []
        for b in self.config.boundaries:
            if b.type == "BOUND_PERIODIC":
                out.append((self.mesh.group_index(b.name), self.mesh.group_index(b.pair),
                            float(b.shift[0]), float(b.shift[1])))
        return out

    def initial_state(self):
        if self.U0 is not None:
            return np.ascontiguousarray(self.U0, dtype=np.float64)
        regs = {r.name: r for r in self.config.regions}
        U = np.empty((4, self.mesh.n_cells))
        for gi, gname in enumerate(self.mesh.cell_group_names):
            if gname not in regs:
                raise KeyError("ERROR: unknown cell name '%s'" % gname)
            r = regs[gname]
            U[:, self.mesh.cell_group == gi] = region_state(r, self.config.materials[r.material])[:, None]
        return U


def shock_tube(nx=200, ny=20, seed=1, right=(1.0, 0.5), mild=True, **cfg):
    """[0,1]x[0,0.2] tube, slip walls top/bottom, outlets at the ends.

    right = (T, P) of the right region; the left region is T = P = 1.  mild=True
    is the (rho, p) = (1, 1) | (0.5, 0.5) state the reference can run without
    a limiter; mild=False is Sod (1, 1) | (0.125, 0.1).
    """
    m = meshgen.rect_mesh(nx, ny, 1.0, 0.2, seed=seed)
    if not mild:
        right = (0.8, 0.1)
    regions = [Region("cells_left", 0.0, 0.0, 1.0, 1.0), Region("cells_right", 0.0, 0.0, right[0], right[1])]
    bnds = [Boundary("edge_bottom", "BOUND_WALL_SLIP"), Boundary("edge_top", "BOUND_WALL_NO_SLIP"),
            Boundary("edge_left", "BOUND_OUTLET"), Boundary("edge_right", "BOUND_OUTLET")]
    c = SolverConfig(regions=regions, boundaries=bnds, **cfg)
    return Case(m, c, name="shock_tube_%dx%d" % (nx, ny))


def periodic_pulse(n=1024, seed=2, mach=0.5, angle_deg=30.0, amp=0.2, sigma=0.05, noise=1.0e-3,
                   mu=1.0e-4, prandtl=0.72, **cfg):
    """Periodic unit square: uniform flow plus a Gaussian density/pressure pulse.

    rho0 = 1, p0 = 1/gamma (c0 = 1), |u| = mach at angle_deg;
    rho = rho0 (1 + amp g), p = p0 (1 + amp g), g = exp(-d^2 / (2 sigma^2)) with
    d the periodic distance from the cell centroid to (0.5, 0.5); seeded
    U(-noise, noise) added to both velocity components of every cell.
    Navier-Stokes with mu and Pr, least-squares gradients, SSP-RK3, CFL 0.4
    with the spectral time step unless overridden.
    """
    m = meshgen.periodic_square(n, seed=seed)
    mat = Material(name="air", ML=mu)
    gamma = mat.gamma
    c = m.centroids()
    dx = np.abs(c[:, 0] - 0.5)
    dy = np.abs(c[:, 1] - 0.5)
    dx = np.minimum(dx, 1.0 - dx)
    dy = np.minimum(dy, 1.0 - dy)
    g = np.exp(-(dx * dx + dy * dy) / (2.0 * sigma * sigma))
    rho = 1.0 + amp * g
    p = (1.0 / gamma) * (1.0 + amp * g)
    ang = math.radians(angle_deg)
    idx = np.arange(m.n_cells)
    u = mach * math.cos(ang) + noise * meshgen.uniform(seed, 40, idx)
    v = mach * math.sin(ang) + noise * meshgen.uniform(seed, 41, idx)
    U0 = prim_to_cons(rho, u, v, p, gamma)
    defaults = dict(flux="GODUNOV", gradient="LSQ", limiter="NONE", rk_stages=3, viscous=True,
                    cfl=0.4, cfl_mode="SPECTRAL", prandtl=prandtl)
    defaults.update(cfg)
    bnds = [Boundary("edge_right", "BOUND_PERIODIC", pair="edge_left", shift=(1.0, 0.0)),
            Boundary("edge_top", "BOUND_PERIODIC", pair="edge_bottom", shift=(0.0, 1.0))]
    conf = SolverConfig(materials=[mat], boundaries=bnds, regions=[Region("fluid", 0, 0, 1, 1)], **defaults)
    return Case(m, conf, U0=U0, name="periodic_pulse_%d" % n)


def inflow_state(mach, angle_deg, gamma=1.4, rho=1.0, p=None):
    """Free-stream (u, v, T, p) for rho = 1, p = 1/gamma (so c = 1) and R = 1."""
    if p is None:
        p = 1.0 / gamma
    ang = math.radians(angle_deg)
    c = math.sqrt(gamma * p / rho)
    return mach * c * math.cos(ang), mach * c * math.sin(ang), p / rho, p


def bump_channel(ny=20, height=0.05, mach=0.8, angle_deg=0.0, seed=3, outlet="BOUND_PRESSURE_OUTLET", **cfg):
    """Channel with a circular-arc bump: inlet (Vx, Vy, T, P) from Mach and angle,
    pressure outlet at the free-stream static pressure (or outlet="BOUND_OUTLET",
    the extrapolating outlet the implicit method knows), slip walls elsewhere;
    the channel starts at the free-stream state."""
    m = meshgen.bump_channel(ny, height, seed=seed)
    mat = Material()
    u, v, T, p = inflow_state(mach, angle_deg, mat.gamma)
    bnds = [Boundary("inlet", "BOUND_INLET", Vx=u, Vy=v, T=T, P=p),
            Boundary("outlet", outlet, P=p),
            Boundary("wall_bottom", "BOUND_WALL_SLIP"), Boundary("bump", "BOUND_WALL_SLIP"),
            Boundary("wall_top", "BOUND_WALL_SLIP")]
    defaults = dict(flux="GODUNOV", gradient="GREEN_GAUSS", limiter="MINMOD", rk_stages=2, cfl=0.5)
    defaults.update(cfg)
    conf = SolverConfig(materials=[mat], boundaries=bnds, regions=[Region("fluid", u, v, T, p)], **defaults)
    return Case(m, conf, name="bump_%d" % ny, meta=dict(mach=mach, angle_deg=angle_deg, height=height))
