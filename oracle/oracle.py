"""ctypes binding of the C oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py -- never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


class OrCfg(C.Structure):
    _fields_ = [("flux", C.c_int), ("gradient", C.c_int), ("limiter", C.c_int), ("rk_stages", C.c_int),
                ("viscous", C.c_int), ("steady", C.c_int), ("cfl_mode", C.c_int), ("max_newton", C.c_int),
                ("cfl", C.c_double), ("tau", C.c_double), ("gam_flux", C.c_double),
                ("ro_min", C.c_double), ("ro_max", C.c_double), ("p_min", C.c_double),
                ("p_max", C.c_double), ("u_max", C.c_double), ("prandtl", C.c_double),
                ("nmat", C.c_int), ("M", C.c_double * 8), ("Cp", C.c_double * 8), ("K", C.c_double * 8),
                ("ML", C.c_double * 8), ("nbc", C.c_int), ("bc_type", C.c_int * 64),
                ("bc_par", (C.c_double * 4) * 64)]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.or_mesh_build.argtypes = [C.c_int, _dp, C.c_int, _ip, C.c_int, _ip, _ip, C.POINTER(vp)]
        L.or_mesh_periodic.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double]
        L.or_mesh_free.argtypes = [vp]
        L.or_mesh_counts.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_mesh_export.argtypes = [vp, _dp, _ip, _ip, _dp, _ip, _ip, _dp]
        L.or_solver_create.restype = vp
        L.or_solver_create.argtypes = [vp, C.POINTER(OrCfg), _ip, _ip]
        L.or_solver_free.argtypes = [vp]
        for n in ("or_set_state", "or_get_state", "or_gradients", "or_edge_fluxes", "or_residual"):
            getattr(L, n).argtypes = [vp, _dp]
        L.or_set_flags.argtypes = [vp, _up]
        L.or_get_flags.argtypes = [vp, _up]
        L.or_calc_tau.argtypes = [vp]
        L.or_get_tau.restype = C.c_double
        L.or_get_tau.argtypes = [vp, _dp]
        L.or_set_tau.argtypes = [vp, C.c_double]
        L.or_run.argtypes = [vp, C.c_int]
        L.or_newton_cap_hits.restype = C.c_long
        L.or_newton_cap_hits.argtypes = [vp]
        L.or_wall_force.restype = C.c_double
        L.or_wall_force.argtypes = [vp, _ip, C.POINTER(C.c_double)]
        L.or_rim.argtypes = [C.c_int, _dp, _dp, C.c_int, _ip]
        L.or_flux.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_double]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_implicit_assemble.restype = C.c_int
        L.or_implicit_assemble.argtypes = [vp, C.c_int, C.c_double, C.c_double, _dp, _ip, _dp, _dp]
        L.or_roe_jacobians.argtypes = [C.c_int, _dp, _dp, C.c_double]
        _lib = L
    return _lib


class OracleMesh:
    def __init__(self, mesh, periodic=()):
        L = lib()
        self._h = C.c_void_p()
        xy = np.ascontiguousarray(mesh.xy, dtype=np.float64)
        tri = np.ascontiguousarray(mesh.tri, dtype=np.int32)
        be = np.ascontiguousarray(mesh.bedges, dtype=np.int32)
        bg = np.ascontiguousarray(mesh.bedge_group, dtype=np.int32)
        if be.size == 0:
            be = np.zeros((1, 2), np.int32)
            bg = np.zeros(1, np.int32)
        rc = L.or_mesh_build(mesh.n_nodes, xy.ravel(), mesh.n_cells, tri.ravel(), mesh.n_bedges,
                             be.ravel(), bg, C.byref(self._h))
        self.build_status = rc
        for (ga, gb, sx, sy) in periodic:
            rc = L.or_mesh_periodic(self._h, ga, gb, sx, sy)
            if rc:
                raise RuntimeError("periodic pairing failed (%d)" % rc)
        nn, nc, ne = C.c_int(), C.c_int(), C.c_int()
        L.or_mesh_counts(self._h, C.byref(nn), C.byref(nc), C.byref(ne))
        self.nn, self.nc, self.ne = nn.value, nc.value, ne.value

    def export(self):
        nc, ne = self.nc, self.ne
        out = dict(cell_geom=np.zeros((nc, 4)), cell_nodes=np.zeros((nc, 3), np.int32),
                   cell_edges=np.zeros((nc, 3), np.int32), edge_geom=np.zeros((ne, 9)),
                   edge_int=np.zeros((ne, 5), np.int32), edge_group=np.zeros(ne, np.int32),
                   edge_shift=np.zeros((ne, 2)))
        lib().or_mesh_export(self._h, out["cell_geom"].ravel(), out["cell_nodes"].ravel(),
                             out["cell_edges"].ravel(), out["edge_geom"].ravel(), out["edge_int"].ravel(),
                             out["edge_group"], out["edge_shift"].ravel())
        return out

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.or_mesh_free(self._h)
            self._h = None


class OracleSolver:
    """The oracle on one case (cfd2d.case.Case)."""

    def __init__(self, case, threads=1):
        L = lib()
        L.or_set_threads(threads)
        self.case = case
        self.mesh = OracleMesh(case.mesh, case.periodic_pairs())
        if self.mesh.build_status:
            raise RuntimeError("mesh build failed (%d)" % self.mesh.build_status)
        ex = self.mesh.export()
        self.geometry = ex
        gbc = case.group_bc()
        eg = ex["edge_group"]
        ebc = np.where(eg >= 0, gbc[np.maximum(eg, 0)], -1).astype(np.int32)
        inner = ex["edge_int"][:, 3] >= 0
        if np.any(~inner & (ebc < 0)):
            raise RuntimeError("ERROR (boundary condition): unknown edge type")
        self.edge_bc = ebc
        self.cfg = case.config.fill(OrCfg())
        self._h = L.or_solver_create(self.mesh._h, C.byref(self.cfg),
                                     np.ascontiguousarray(case.cell_material(), np.int32), ebc)
        self.nc, self.ne = self.mesh.nc, self.mesh.ne
        self.set_state(case.initial_state())
        L.or_calc_tau(self._h)

    def set_state(self, U):
        lib().or_set_state(self._h, np.ascontiguousarray(U, dtype=np.float64).ravel())

    def state(self):
        U = np.zeros(4 * self.nc)
        lib().or_get_state(self._h, U)
        return U.reshape(4, self.nc)

    def flags(self):
        f = np.zeros(self.nc, np.uint32)
        lib().or_get_flags(self._h, f)
        return f

    def set_flags(self, f):
        lib().or_set_flags(self._h, np.ascontiguousarray(f, dtype=np.uint32))

    def tau(self):
        ct = np.zeros(self.nc)
        t = lib().or_get_tau(self._h, ct)
        return t, ct

    def set_tau(self, tau):
        lib().or_set_tau(self._h, tau)

    def gradients(self):
        g = np.zeros(8 * self.nc)
        lib().or_gradients(self._h, g)
        return g.reshape(8, self.nc)

    def edge_fluxes(self):
        f = np.zeros(4 * self.ne)
        lib().or_edge_fluxes(self._h, f)
        return f.reshape(self.ne, 4)

    def residual(self):
        r = np.zeros(4 * self.nc)
        lib().or_residual(self._h, r)
        return r.reshape(4, self.nc)

    def run(self, n):
        lib().or_run(self._h, int(n))

    def newton_cap_hits(self):
        return lib().or_newton_cap_hits(self._h)

    def wall_force(self, edge_mask):
        fx = C.c_double()
        fy = lib().or_wall_force(self._h, np.ascontiguousarray(edge_mask, np.int32), C.byref(fx))
        return fx.value, fy

    def implicit_assemble(self, steady=False, cfl=0.5, tau=1.0e-3):
        """The implicit method's block rows: (blocks[nc,4,4,4], cols[nc,4], rhs[nc,4], ctau[nc])."""
        nc = self.nc
        blocks = np.zeros(64 * nc)
        cols = np.zeros(4 * nc, np.int32)
        rhs = np.zeros(4 * nc)
        ctau = np.zeros(nc)
        rc = lib().or_implicit_assemble(self._h, int(bool(steady)), float(cfl), float(tau), blocks, cols, rhs, ctau)
        if rc:
            raise ValueError("implicit assembly: %s" % ("pressure outlet present" if rc == 1 else "flux must be GODUNOV or LAX"))
        return blocks.reshape(nc, 4, 4, 4), cols.reshape(nc, 4), rhs.reshape(nc, 4), ctau

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.or_solver_free(self._h)
            self._h = None


def rim(rows, max_iter=0):
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    n = rows.shape[0]
    out = np.zeros((n, 6))
    it = np.zeros(n, np.int32)
    lib().or_rim(n, rows.ravel(), out.ravel(), max_iter, it)
    return out, it


def flux(kind, rows, gam=1.4):
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    out = np.zeros((rows.shape[0], 4))
    lib().or_flux(kind, rows.shape[0], rows.ravel(), out.ravel(), gam)
    return out


def roe_jacobians(rows, gam=1.4):
    """Rows (rL,pL,uL,vL,rR,pR,uR,vR,nx,ny) -> (roe[n,6] = r,u,v,p,c,H; A+[n,4,4]; A-[n,4,4]) for l = 1."""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    n = rows.shape[0]
    out = np.zeros((n, 38))
    lib().or_roe_jacobians(n, rows.ravel(), out.ravel(), gam)
    return out[:, :6], out[:, 6:22].reshape(n, 4, 4), out[:, 22:38].reshape(n, 4, 4)
