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
        self.bc_off = [0]
        for c, part in zip(self.cases, parts):
            base = len(bcs)
            own = [b for b in c.config.boundaries if b.type != "BOUND_PERIODIC"]
            bcs.extend(own)
            self.bc_off.append(len(bcs))
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

    def autotune(self, trial_steps=3):
        """Picks the fastest step variant for this batch (state restored)."""
        return self.ctx.autotune(trial_steps)

    def run(self, n, graph=False):
        if graph:
            self.ctx.step_graph(n)
        else:
            self.ctx.step(n)
        self.ctx.sync()

    def states(self):
        U = self.ctx.state()
        return [U[:, self.cell_off[k]:self.cell_off[k + 1]] for k in range(len(self.cases))]

    def edge_mask(self, group):
        """int32 mask of the edges in the named edge group of every case (the
        concatenated mesh names case k's groups "k:name")."""
        names = self.mesh.group_names(0)
        ids = [i for i, n in enumerate(names) if n.split(":", 1)[-1] == group]
        eg = self.mesh.export()["edge_group"]
        return np.isin(eg, ids).astype(np.int32)

    def adjoint(self, nsteps, edge_mask, ckpt_every=0):
        """One reverse sweep for every case together: the cases are
        independent, so the adjoint of the summed objective gives each case's
        own gradients.  Returns per case (J_k, dJ_k/dU0 (4, nc_k),
        dJ_k/dpar (n_bc_k, 4)); the context ends at the final state."""
        _, gU, gb = self.ctx.adjoint(nsteps, edge_mask, ckpt_every)
        Jc = self.ctx.adjoint_objectives(len(self.cases))
        return [(Jc[k], gU[:, self.cell_off[k]:self.cell_off[k + 1]], gb[self.bc_off[k]:self.bc_off[k + 1]])
                for k in range(len(self.cases))]

    def adjoint_geometry(self, nsteps, edge_mask, ckpt_every=0):
        """adjoint() plus each case's dJ_k/d(its node coordinates) as
        (n_nodes_k, 2) (cfd2d_adjoint_geometry on the concatenated mesh, whose
        nodes are the cases' nodes in order)."""
        _, gU, gb, gx = self.ctx.adjoint_geometry(nsteps, edge_mask, ckpt_every)
        Jc = self.ctx.adjoint_objectives(len(self.cases))
        no = np.concatenate([[0], np.cumsum([c.mesh.n_nodes for c in self.cases])])
        return [(Jc[k], gU[:, self.cell_off[k]:self.cell_off[k + 1]], gb[self.bc_off[k]:self.bc_off[k + 1]],
                 gx[no[k]:no[k + 1]]) for k in range(len(self.cases))]

    def wall_force_y(self, edge_mask):
        """Per case: sum over its masked edges of p(c1) * l * n_y on the current
        state (one download of the state for all cases)."""
        if not hasattr(self, "_edge_geom"):
            ex = self.mesh.export()
            self._edge_geom = (ex["edge_int"][:, 2].copy(), ex["edge_geom"][:, 2] * ex["edge_geom"][:, 1])
        c1, lny = self._edge_geom
        U = self.ctx.state()
        gam = self.cases[0].config.materials[0].gamma
        r = U[0]
        u, v, E = U[1] / r, U[2] / r, U[3] / r
        p = r * (E - 0.5 * (u * u + v * v)) * (gam - 1)
        out = []
        for k in range(len(self.cases)):
            e = np.nonzero(edge_mask[self.edge_off[k]:self.edge_off[k + 1]])[0] + self.edge_off[k]
            out.append(float(np.sum(p[c1[e]] * lny[e])))
        return out

    def close(self):
        self.ctx.close()
        self.mesh.close()
