"""Many independent cases advanced together in one index space.

The meshes of all cases are concatenated (cfd2d_mesh_concat); every kernel
launch of a step then covers the cells/edges of all cases at once, each
case keeps its own boundary objects and its own time step (TAU per case,
fvm_tvd.cpp:226-237 applied case by case).  The numerical method, its
options and the materials must be the same for all cases.
"""
from __future__ import annotations

import numpy as np

from . import native


class BatchSolver:
    def __init__(self, cases, calc_tau=True):
        if not cases:
            raise ValueError("no cases")
        self.cases = list(cases)
        c0 = self.cases[0].config
        for c in self.cases[1:]:
            k = c.config
            same = (k.flux, k.gradient, k.limiter, k.rk_stages, k.viscous, k.steady, k.cfl_mode, k.cfl,
                    k.gam_flux) == (c0.flux, c0.gradient, c0.limiter, c0.rk_stages, c0.viscous, c0.steady,
                                    c0.cfl_mode, c0.cfl, c0.gam_flux)
            if not same or len(k.materials) != len(c0.materials):
                raise ValueError("all batched cases need the same method options and materials")
        parts = [native.Mesh.from_trimesh(c.mesh, c.periodic_pairs()) for c in self.cases]
        self.mesh = native.Mesh.concat(parts)
        self.cell_off, self.edge_off = self.mesh.case_offsets()
        # boundary objects of all cases, groups renumbered per part
        bcs, gbc, cmat, U = [], [], [], []
        for c, part in zip(self.cases, parts):
            base = len(bcs)
            own = [b for b in c.config.boundaries if b.type != "BOUND_PERIODIC"]
            bcs.extend(own)
            g = c.group_bc()
            gbc.extend([base + v if v >= 0 else -1 for v in g])
            cmat.append(c.cell_material())
            U.append(c.initial_state())
            part.close()
        cfg = type(c0)(**{**c0.__dict__, "boundaries": bcs})
        self.ctx = native.Context(self.mesh, cfg, np.asarray(gbc, np.int32), np.concatenate(cmat))
        self.nc, self.ne = self.mesh.nc, self.mesh.ne
        self.ctx.set_state(np.concatenate(U, axis=1))
        if calc_tau:
            self.ctx.calc_tau()

    def run(self, n, graph=False):
        if graph:
            self.ctx.step_graph(n)
        else:
            self.ctx.step(n)
        self.ctx.sync()

    def states(self):
        U = self.ctx.state()
        return [U[:, self.cell_off[k]:self.cell_off[k + 1]] for k in range(len(self.cases))]

    def close(self):
        self.ctx.close()
        self.mesh.close()
