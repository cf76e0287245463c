"""The implicit finite-volume method with its matrix fill on the GPU.

FVM_TVD_IMPLICIT::run (src/methods/fvm_tvd_implicit.cpp:478-738) per step:
assemble the block system [S/tau I + sum A+ l] dU_i + sum A- l dU_j = RHS_i
(here on the device: cfd2d_implicit_assemble), solve it on the host
(HYPRE_GMRES in the reference: GMRES(30), relative tolerance 1e-7, at most
20 iterations, zero start, SolverHypreGmres.cpp:8-98; here the restarted
GMRES below), then apply the increment to the cells that are not LIM-flagged,
clamp to the limits (clamped cells get the clamped state and the LIM flag,
671-692), average LIM cells over their neighbours (remediateLimCells,
736-768), and in steady runs lower the CFL when too many cells are limited
or the solve fails (decCFL 1070-1077) and raise it every cfl_step steps
(incCFL 1079-1086).

The state lives on the host (the solver's side) and is handed to the
device before each assembly.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp

from . import native

CELL_FLAG_LIM = 0x2  # src/mesh/grid.h:14
RESULT_OK, RESULT_ERR_MAX_ITER, RESULT_ERR_CONVERG = 0x0, 0x2, 0x4  # MatrixSolver.h:13-16


def bsr_from_rows(blocks, cols):
    """scipy BSR matrix (4x4 blocks) from the block rows of cfd2d_implicit_assemble
    (slots with column -1 dropped, blocks of repeated columns summed)."""
    nc = blocks.shape[0]
    valid = cols >= 0
    indptr = np.zeros(nc + 1, np.int64)
    np.cumsum(valid.sum(axis=1), out=indptr[1:])
    A = sp.bsr_matrix((blocks[valid], cols[valid], indptr), shape=(4 * nc, 4 * nc))
    A.sum_duplicates()
    return A


def gmres(A, b, tol=1.0e-7, max_iter=20, kdim=30):
    """Restarted GMRES (Saad, Iterative Methods for Sparse Linear Systems,
    2nd ed., alg. 6.11, Givens rotations) from x = 0 with the stopping rule
    of hypre's GMRES: relative residual |r| / |b| < tol, at most max_iter
    iterations in total.  Returns (x, iterations, final relative residual)."""
    n = b.size
    x = np.zeros(n)
    bn = float(np.linalg.norm(b))
    if bn == 0.0:
        return x, 0, 0.0
    it = 0
    rel = 1.0
    while it < max_iter:
        r = b - A @ x
        beta = float(np.linalg.norm(r))
        rel = beta / bn
        if rel < tol:
            break
        m = min(kdim, max_iter - it)
        V = np.zeros((m + 1, n))
        H = np.zeros((m + 1, m))
        cs = np.zeros(m)
        sn = np.zeros(m)
        g = np.zeros(m + 1)
        g[0] = beta
        V[0] = r / beta
        k = 0
        for j in range(m):
            w = A @ V[j]
            for i in range(j + 1):  # modified Gram-Schmidt
                H[i, j] = float(w @ V[i])
                w = w - H[i, j] * V[i]
            H[j + 1, j] = float(np.linalg.norm(w))
            if H[j + 1, j] != 0.0:
                V[j + 1] = w / H[j + 1, j]
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            d = math.hypot(H[j, j], H[j + 1, j])
            cs[j], sn[j] = (1.0, 0.0) if d == 0.0 else (H[j, j] / d, H[j + 1, j] / d)
            H[j, j] = d
            H[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            it += 1
            k = j + 1
            rel = abs(g[j + 1]) / bn
            if rel < tol or H[j, j] == 0.0:
                break
        y = np.zeros(k)
        for i in range(k - 1, -1, -1):
            y[i] = (g[i] - H[i, i + 1:k] @ y[i + 1:k]) / H[i, i] if H[i, i] != 0.0 else 0.0
        x = x + y @ V[:k]
        if rel < tol:
            break
    rel = float(np.linalg.norm(b - A @ x)) / bn
    return x, it, rel


class ImplicitSolver:
    """One case advanced by the implicit method.

    steady: local time step 2 CFL S / sum(l (|u|+c)) with the CFL ramp
    (cfl start, cfl_scale, cfl_max, cfl_step, max_lim_cells); otherwise the
    global time step tau (the task's TAU).  smoothing > 0 applies the
    increment smoothing of smoothingDelta (fvm_tvd_implicit.cpp:790-808) as
    evidently intended: the reference indexes its per-cell sums with the
    flattened unknown index 4i+k, past their length (undefined behaviour);
    here they are indexed by cell.

    assemble(U, steady, cfl, tau) -> (blocks, cols, rhs, ctau) may replace
    the device assembly (the default) with another implementation of the
    same block rows.
    """

    def __init__(self, case, steady=None, cfl=None, cfl_scale=1.0, cfl_max=None, cfl_step=1, max_lim_cells=None,
                 tau=None, smoothing=0.0, eps=1.0e-7, max_iter=20, kdim=30, assemble=None):
        cfg = case.config
        self.case = case
        self.steady = bool(cfg.steady if steady is None else steady)
        self.cfl = float(cfg.cfl if cfl is None else cfl)
        self.cfl_scale = float(cfl_scale)
        self.cfl_max = float(self.cfl if cfl_max is None else cfl_max)
        self.cfl_step = max(1, int(cfl_step))
        self.max_lim_cells = max_lim_cells
        self.tau = float(cfg.tau if tau is None else tau)
        self.smoothing = float(smoothing)
        self.eps, self.max_iter, self.kdim = eps, max_iter, kdim
        self.limits = (cfg.ro_min, cfg.ro_max, cfg.p_min, cfg.p_max, cfg.u_max)
        self.gam = np.array([m.gamma for m in cfg.materials])[case.cell_material()]
        self.U = np.array(case.initial_state(), dtype=np.float64, copy=True)
        self.nc = self.U.shape[1]
        self.flags = np.zeros(self.nc, np.uint32)
        self.t = 0.0
        self.step_count = 0
        self.solver_err = 0
        self.last_iters = 0
        self.last_residual = 0.0
        self.gpu = None
        self._bufs = None
        if assemble is None:
            self.gpu = native.GpuSolver(case, calc_tau=False)
            assemble = self._gpu_assemble
        self._assemble = assemble
        mesh = case.mesh
        self._nbrs = None
        self._S = None
        self._mesh = mesh

    def _gpu_assemble(self, U, steady, cfl, tau):
        self.gpu.set_state(U)
        if self._bufs is None:
            self._bufs = self.gpu.implicit_buffers()  # page-locked, reused every step
        return self.gpu.implicit_assemble(steady=steady, cfl=cfl, tau=tau, out=self._bufs)

    def _topology(self):
        if self._nbrs is None:
            m = native.Mesh.from_trimesh(self._mesh, self.case.periodic_pairs())
            ex = m.export()
            m.close()
            ei = ex["edge_int"]
            ce = ex["cell_edges"]
            c1, c2 = ei[ce, 2], ei[ce, 3]
            own = np.arange(self.nc)[:, None]
            self._nbrs = np.where(c2 == own, c1, c2)  # neighbour across each face, -1 on a boundary
            self._S = ex["cell_geom"][:, 0].copy()
            self._inner = ei[:, 3] >= 0
            self._ec = ei[:, 2:4]
        return self._nbrs

    def _smooth(self, x):
        """smoothingDelta with per-cell sums (see the class note)."""
        self._topology()
        lim = (self.flags & CELL_FLAG_LIM) > 0
        c1, c2 = self._ec[self._inner, 0], self._ec[self._inner, 1]
        X = x.reshape(self.nc, 4)
        a1 = ~lim[c2]  # c1 takes the increment of a live c2, and the other way round
        a2 = ~lim[c1]
        cnt = np.bincount(c1[a1], minlength=self.nc) + np.bincount(c2[a2], minlength=self.nc)
        for k in range(4):
            acc = np.bincount(c1[a1], weights=X[c2[a1], k], minlength=self.nc) + \
                np.bincount(c2[a2], weights=X[c1[a2], k], minlength=self.nc)
            X[:, k] = (X[:, k] + self.smoothing * acc) / (1.0 + self.smoothing * cnt)
        return X.reshape(-1)

    def _apply(self, x):
        """fvm_tvd_implicit.cpp:675-692: increment, limits, LIM flags."""
        U = self.U
        X = x.reshape(self.nc, 4)
        live = (self.flags & CELL_FLAG_LIM) == 0
        U[:, live] += X[live].T
        ro_min, ro_max, p_min, p_max, u_max = self.limits
        r = U[0].copy()
        u = U[1] / U[0]
        v = U[2] / U[0]
        e = U[3] / U[0] - 0.5 * (u * u + v * v)
        p = r * e * (self.gam - 1.0)
        hit = np.zeros(self.nc, bool)

        def clamp(arr, cond, val):
            nonlocal hit
            c = live & cond
            if c.any():
                arr[c] = val(arr[c]) if callable(val) else val
                hit |= c

        clamp(r, r > ro_max, ro_max)
        clamp(r, r < ro_min, ro_min)
        clamp(u, np.abs(u) > u_max, lambda a: u_max * a / np.abs(a))
        clamp(v, np.abs(v) > u_max, lambda a: u_max * a / np.abs(a))
        clamp(p, p > p_max, p_max)
        clamp(p, p < p_min, p_min)
        self.flags[hit] |= CELL_FLAG_LIM
        if hit.any():
            e2 = p[hit] / ((1.4 - 1.0) * r[hit])  # the reference's __GAM here
            U[0, hit] = r[hit]
            U[1, hit] = r[hit] * u[hit]
            U[2, hit] = r[hit] * v[hit]
            U[3, hit] = r[hit] * (e2 + 0.5 * (u[hit] ** 2 + v[hit] ** 2))

    def _remediate(self):
        """remediateLimCells (fvm_tvd_implicit.cpp:736-768), in cell order, in place."""
        lim = np.nonzero(self.flags & CELL_FLAG_LIM)[0]
        if lim.size == 0:
            return
        nb = self._topology()
        S = self._S
        for i in lim:
            js = nb[i][nb[i] >= 0]
            Ssum = 0.0
            acc = np.zeros(4)
            for j in js:
                Ssum += S[j]
                acc += self.U[:, j] * S[j]
            if Ssum >= self.tau * self.tau:
                self.U[:, i] = acc / Ssum
            f = int(self.flags[i]) + 0x010000
            if f & 0x200000:
                f &= 0x001110
            self.flags[i] = f

    def _dec_cfl(self):
        if self.cfl > 0.1:
            self.cfl = max(0.1, self.cfl * 0.75)

    def _inc_cfl(self):
        if self.cfl < self.cfl_max:
            self.cfl = min(self.cfl_max, self.cfl * self.cfl_scale)

    def step(self):
        """One pass of the run loop; returns the solver result code."""
        if not self.steady:
            self.t += self.tau
        if not self.solver_err:
            self.step_count += 1
        blocks, cols, rhs, ctau = self._assemble(self.U, self.steady, self.cfl, self.tau)
        self.ctau = ctau
        A = bsr_from_rows(blocks, cols)
        x, iters, rel = gmres(A, rhs.reshape(-1), self.eps, self.max_iter, self.kdim)
        self.last_iters, self.last_residual = iters, rel
        res = RESULT_OK
        if iters >= self.max_iter:
            res |= RESULT_ERR_MAX_ITER
        if not (rel < self.eps) or not np.isfinite(rel):
            res |= RESULT_ERR_CONVERG
        self.solver_err = res
        if res == RESULT_OK:
            if self.smoothing > 0.0:
                x = self._smooth(x)
            self._apply(x)
            self._remediate()
            nlim = int(((self.flags & CELL_FLAG_LIM) > 0).sum())
            if self.steady and self.max_lim_cells is not None and nlim >= self.max_lim_cells:
                self._dec_cfl()
            if self.steady and self.step_count % self.cfl_step == 0:
                self._inc_cfl()
        elif self.steady:
            self._dec_cfl()
        else:
            self.solver_err = 0
        return res

    def run(self, nsteps, tmax=math.inf, max_passes=None):
        """Steps while t < tmax and fewer than nsteps steps have been taken.
        A failed steady solve repeats the step at a lower CFL, as the
        reference does; max_passes (default 4 nsteps + 10) bounds the passes
        and raises RuntimeError when the solver keeps failing."""
        target = self.step_count + nsteps
        limit = 4 * nsteps + 10 if max_passes is None else max_passes
        passes = 0
        while self.t < tmax and self.step_count < target:
            if passes >= limit:
                raise RuntimeError("implicit step %d: linear solver failed %d times (last relative residual %.3e)"
                                   % (self.step_count, passes, self.last_residual))
            self.step()
            passes += 1
        return self

    def state(self):
        return self.U.copy()

    def close(self):
        self._bufs = None
        if self.gpu is not None:
            self.gpu.close()
            self.gpu = None
