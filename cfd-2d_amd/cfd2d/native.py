"""ctypes binding of libcfd2d.so (include/cfd2d.h).

The library is built in-tree by cfd-2d_amd/csrc/Makefile.  There is no
fallback: if the library or a HIP device is missing, the calls fail loudly.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CFD2D_LIB") or os.path.join(HERE, "libcfd2d.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")

K_GRADIENT, K_EDGE_FLUX, K_UPDATE, K_REMEDIATE, K_TAU, K_PRIM = range(6)
KERNEL_NAMES = ["gradient", "edge_flux", "update", "remediate", "tau", "prim", "impl_gradient", "impl_edge",
                "impl_assemble"]


class Params(C.Structure):
    _fields_ = [("flux", C.c_int), ("gradient", C.c_int), ("limiter", C.c_int), ("rk_stages", C.c_int),
                ("viscous", C.c_int), ("steady", C.c_int), ("cfl_mode", C.c_int), ("max_newton", C.c_int),
                ("cfl", C.c_double), ("tau", C.c_double), ("gam_flux", C.c_double),
                ("ro_min", C.c_double), ("ro_max", C.c_double), ("p_min", C.c_double),
                ("p_max", C.c_double), ("u_max", C.c_double), ("prandtl", C.c_double),
                ("nmat", C.c_int), ("M", C.c_double * 8), ("Cp", C.c_double * 8), ("K", C.c_double * 8),
                ("ML", C.c_double * 8)]


class Cfd2dError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("cfd2d error %d: %s" % (code, msg))
        self.code = code


# every exported symbol of include/cfd2d.h with its argument types
SIGNATURES = {
    "cfd2d_error_string": (C.c_char_p, [C.c_int]),
    "cfd2d_last_error": (C.c_char_p, []),
    "cfd2d_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "cfd2d_mesh_create": (C.c_int, [C.c_int, _dp, C.c_int, _ip, C.c_int, _ip, _ip, C.POINTER(C.c_void_p)]),
    "cfd2d_mesh_read": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    "cfd2d_mesh_set_groups": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_char_p)]),
    "cfd2d_mesh_set_cell_groups": (C.c_int, [C.c_void_p, _ip]),
    "cfd2d_mesh_periodic": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double]),
    "cfd2d_mesh_concat": (C.c_int, [C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "cfd2d_mesh_renumber": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]),
    "cfd2d_mesh_counts": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "cfd2d_mesh_nodes": (C.c_int, [C.c_void_p, _dp]),
    "cfd2d_mesh_group_count": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "cfd2d_mesh_group_name": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.c_int]),
    "cfd2d_mesh_get_cell_groups": (C.c_int, [C.c_void_p, _ip]),
    "cfd2d_mesh_case_offsets": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_void_p, C.c_void_p]),
    "cfd2d_mesh_export": (C.c_int, [C.c_void_p, _dp, _ip, _ip, _dp, _ip, _ip, _dp]),
    "cfd2d_mesh_destroy": (None, [C.c_void_p]),
    "cfd2d_mesh_geometry_vjp": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, _dp]),
    "cfd2d_ctx_create": (C.c_int, [C.c_void_p, C.POINTER(Params), C.c_int, _ip, _dp, _ip, _ip,
                                   C.POINTER(C.c_void_p)]),
    "cfd2d_ctx_destroy": (None, [C.c_void_p]),
    "cfd2d_set_state": (C.c_int, [C.c_void_p, _dp]),
    "cfd2d_get_state": (C.c_int, [C.c_void_p, _dp]),
    "cfd2d_set_state_device": (C.c_int, [C.c_void_p, C.c_void_p]),
    "cfd2d_get_state_device": (C.c_int, [C.c_void_p, C.c_void_p]),
    "cfd2d_set_flags": (C.c_int, [C.c_void_p, _up]),
    "cfd2d_get_flags": (C.c_int, [C.c_void_p, _up]),
    "cfd2d_calc_tau": (C.c_int, [C.c_void_p]),
    "cfd2d_get_tau": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_void_p]),
    "cfd2d_set_tau": (C.c_int, [C.c_void_p, C.c_double]),
    "cfd2d_set_bc_par": (C.c_int, [C.c_void_p, _dp]),
    "cfd2d_gradients": (C.c_int, [C.c_void_p, _dp]),
    "cfd2d_edge_fluxes": (C.c_int, [C.c_void_p, _dp]),
    "cfd2d_residual": (C.c_int, [C.c_void_p, _dp]),
    "cfd2d_step": (C.c_int, [C.c_void_p, C.c_int]),
    "cfd2d_step_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "cfd2d_sync": (C.c_int, [C.c_void_p]),
    "cfd2d_status": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_int)]),
    "cfd2d_wall_force": (C.c_int, [C.c_void_p, _ip, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "cfd2d_adjoint": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _ip, C.POINTER(C.c_double), C.c_void_p,
                                C.c_void_p]),
    "cfd2d_adjoint_objectives": (C.c_int, [C.c_void_p, _dp]),
    "cfd2d_adjoint_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "cfd2d_adjoint_geometry": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, _ip, C.POINTER(C.c_double),
                                         C.c_void_p, C.c_void_p, _dp]),
    "cfd2d_geometry_vjp_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, _dp, _dp, C.c_void_p, _dp]),
    "cfd2d_implicit_assemble": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p]),
    "cfd2d_riemann": (C.c_int, [C.c_int, _dp, C.c_double, C.c_int, _dp, C.c_void_p, C.POINTER(C.c_longlong)]),
    "cfd2d_flux": (C.c_int, [C.c_int, C.c_int, _dp, C.c_double, C.c_int, _dp, C.POINTER(C.c_longlong)]),
    "cfd2d_flux_vjp": (C.c_int, [C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_int, _dp]),
    "cfd2d_flux_noalloc": (C.c_int, [C.c_int, C.c_int, _dp, C.c_double, C.c_int, _dp, C.POINTER(C.c_longlong)]),
    "cfd2d_host_alloc": (C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    "cfd2d_host_free": (C.c_int, [C.c_void_p]),
    "cfd2d_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "cfd2d_timing_kernels": (C.c_int, [C.c_void_p, C.c_uint]),
    "cfd2d_fuse_gradient": (C.c_int, [C.c_void_p, C.c_int]),
    "cfd2d_stream_state": (C.c_int, [C.c_void_p, C.c_int]),
    "cfd2d_autotune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "cfd2d_timing_read": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
    "cfd2d_edge_shape": (C.c_int, [C.c_void_p, C.c_int]),
    "cfd2d_tuning": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                               C.POINTER(C.c_int)]),
    "cfd2d_ctx_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "cfd2d_ctx_stream": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
}

_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libcfd2d.so is not built (run __graft_entry__.build() or make -C cfd-2d_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().cfd2d_last_error().decode(errors="replace")
        raise Cfd2dError(rc, msg)


def device_count():
    n = C.c_int(0)
    rc = lib().cfd2d_device_count(C.byref(n))
    return n.value if rc == 0 else 0


def riemann(rows, gam=1.4, max_newton=0):
    """rim_orig on the device for rows (n, 8) = (rho, p, u_n, u_t) left then
    right; returns (out (n, 5) = RI, EI, PI, UI, VI; Newton iterations (n,);
    rows that hit the cap)."""
    rows = np.ascontiguousarray(rows, np.float64).reshape(-1, 8)
    n = rows.shape[0]
    out = np.zeros((n, 5))
    its = np.zeros(n, np.int32)
    hits = C.c_longlong()
    check(lib().cfd2d_riemann(n, rows.ravel(), float(gam), int(max_newton), out.ravel(),
                              its.ctypes.data_as(C.c_void_p), C.byref(hits)))
    return out, its, hits.value


FLUX_KINDS = {"GODUNOV": 0, "LAX": 1, "HLLC": 2, "ROE": 3}
NOALLOC_MAX_ROWS = 65536  # CFD2D_NOALLOC_MAX_ROWS


def flux(kind, rows, gam=1.4, max_newton=0, noalloc=False):
    """A numerical flux on the device for rows (n, 10) = (rL, pL, uL, vL, rR,
    pR, uR, vR, nx, ny); returns ((n, 4) fluxes, rows that hit the Newton cap).
    noalloc=True calls cfd2d_flux_noalloc (no device allocation, at most
    CFD2D_NOALLOC_MAX_ROWS rows)."""
    k = FLUX_KINDS[kind] if isinstance(kind, str) else int(kind)
    rows = np.ascontiguousarray(rows, np.float64).reshape(-1, 10)
    out = np.zeros((rows.shape[0], 4))
    hits = C.c_longlong()
    f = lib().cfd2d_flux_noalloc if noalloc else lib().cfd2d_flux
    check(f(k, rows.shape[0], rows.ravel(), float(gam), int(max_newton), out.ravel(), C.byref(hits)))
    return out, hits.value


def flux_vjp(kind, rows, fbar, gam=1.4, max_newton=0):
    """The vector-Jacobian product of an adjoint flux (GODUNOV, LAX, HLLC) on
    the device for rows (n, 10) as flux() and seeds fbar (n, 4); returns (n,
    10) = adjoints of (rL, pL, uL, vL, rR, pR, uR, vR, nx, ny)."""
    k = FLUX_KINDS[kind] if isinstance(kind, str) else int(kind)
    rows = np.ascontiguousarray(rows, np.float64).reshape(-1, 10)
    fbar = np.ascontiguousarray(fbar, np.float64).reshape(-1, 4)
    if fbar.shape[0] != rows.shape[0]:
        raise ValueError("flux_vjp: one seed row per state row")
    out = np.zeros((rows.shape[0], 10))
    check(lib().cfd2d_flux_vjp(k, rows.shape[0], rows.ravel(), fbar.ravel(), float(gam), int(max_newton),
                               out.ravel()))
    return out


class Mesh:
    """A host mesh (cfd2d_mesh) built from a TriMesh or read from files."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        nn, nc, ne = C.c_int(), C.c_int(), C.c_int()
        check(lib().cfd2d_mesh_counts(self._h, C.byref(nn), C.byref(nc), C.byref(ne)))
        self.nn, self.nc, self.ne = nn.value, nc.value, ne.value

    @classmethod
    def from_trimesh(cls, m, periodic=()):
        L = lib()
        h = C.c_void_p()
        xy = np.ascontiguousarray(m.xy, np.float64).ravel()
        tri = np.ascontiguousarray(m.tri, np.int32).ravel()
        be = np.ascontiguousarray(m.bedges, np.int32).ravel()
        bg = np.ascontiguousarray(m.bedge_group, np.int32)
        if be.size == 0:
            be = np.zeros(2, np.int32)
            bg = np.zeros(1, np.int32)
        check(L.cfd2d_mesh_create(m.n_nodes, xy, m.n_cells, tri, m.n_bedges, be, bg, C.byref(h)))
        names = (C.c_char_p * len(m.edge_group_names))(*[s.encode() for s in m.edge_group_names])
        check(L.cfd2d_mesh_set_groups(h, 0, len(m.edge_group_names), names))
        cnames = (C.c_char_p * len(m.cell_group_names))(*[s.encode() for s in m.cell_group_names])
        check(L.cfd2d_mesh_set_groups(h, 1, len(m.cell_group_names), cnames))
        check(L.cfd2d_mesh_set_cell_groups(h, np.ascontiguousarray(m.cell_group, np.int32)))
        for (ga, gb, sx, sy) in periodic:
            check(L.cfd2d_mesh_periodic(h, ga, gb, sx, sy))
        return cls(h)

    @classmethod
    def read(cls, files_type, name):
        h = C.c_void_p()
        check(lib().cfd2d_mesh_read(files_type.encode(), name.encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def concat(cls, meshes):
        arr = (C.c_void_p * len(meshes))(*[m._h.value for m in meshes])
        h = C.c_void_p()
        check(lib().cfd2d_mesh_concat(len(meshes), arr, C.byref(h)))
        return cls(h)

    ORDERS = {"none": 0, "hilbert": 1, "shuffle": 2}

    def renumber(self, order="hilbert", seed=0):
        """(the mesh with its cells in another order, permutation p: new cell
        k = old cell p[k]); cfd2d_mesh_renumber."""
        p = np.zeros(self.nc, np.int32)
        h = C.c_void_p()
        check(lib().cfd2d_mesh_renumber(self._h, self.ORDERS[order], int(seed), p.ctypes.data_as(C.c_void_p),
                                        C.byref(h)))
        return Mesh(h), p

    def group_names(self, kind=0):
        n = C.c_int()
        check(lib().cfd2d_mesh_group_count(self._h, kind, C.byref(n)))
        out = []
        buf = C.create_string_buffer(256)
        for i in range(n.value):
            check(lib().cfd2d_mesh_group_name(self._h, kind, i, buf, 256))
            out.append(buf.value.decode())
        return out

    def cell_groups(self):
        g = np.zeros(self.nc, np.int32)
        check(lib().cfd2d_mesh_get_cell_groups(self._h, g))
        return g

    def nodes(self):
        xy = np.zeros(2 * self.nn)
        check(lib().cfd2d_mesh_nodes(self._h, xy))
        return xy.reshape(self.nn, 2)

    def case_offsets(self):
        n = C.c_int()
        check(lib().cfd2d_mesh_case_offsets(self._h, C.byref(n), None, None))
        co = np.zeros(n.value + 1, np.int32)
        eo = np.zeros(n.value + 1, np.int32)
        check(lib().cfd2d_mesh_case_offsets(self._h, C.byref(n), co.ctypes.data_as(C.c_void_p),
                                            eo.ctypes.data_as(C.c_void_p)))
        return co, eo

    def export(self):
        nc, ne = self.nc, self.ne
        out = dict(cell_geom=np.zeros((nc, 4)), cell_nodes=np.zeros((nc, 3), np.int32),
                   cell_edges=np.zeros((nc, 3), np.int32), edge_geom=np.zeros((ne, 9)),
                   edge_int=np.zeros((ne, 5), np.int32), edge_group=np.zeros(ne, np.int32),
                   edge_shift=np.zeros((ne, 2)))
        check(lib().cfd2d_mesh_export(self._h, out["cell_geom"].ravel(), out["cell_nodes"].ravel(),
                                      out["cell_edges"].ravel(), out["edge_geom"].ravel(),
                                      out["edge_int"].ravel(), out["edge_group"], out["edge_shift"].ravel()))
        return out

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value and _lib is not None:
            _lib.cfd2d_mesh_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()


class Context:
    """A device context (cfd2d_ctx): state, passes and time steps on the GPU."""

    def __init__(self, mesh: Mesh, config, group_bc, cell_mat=None):
        L = lib()
        self.mesh = mesh
        p = Params()
        config.fill(p)
        bcs = [b for b in config.boundaries if b.type != "BOUND_PERIODIC"]
        bt = np.array([b.code for b in bcs] or [0], np.int32)
        bp = np.array([list(b.params()) for b in bcs] or [[0, 0, 0, 0]], np.float64).ravel()
        gbc = np.ascontiguousarray(group_bc if len(group_bc) else [-1], np.int32)
        cm = np.ascontiguousarray(cell_mat if cell_mat is not None else np.zeros(mesh.nc), np.int32)
        h = C.c_void_p()
        check(L.cfd2d_ctx_create(mesh._h, C.byref(p), len(bcs), bt, bp, gbc, cm, C.byref(h)))
        self._h = h
        self.nc, self.ne = mesh.nc, mesh.ne
        self.n_bc = len(bcs)

    def set_state(self, U):
        check(lib().cfd2d_set_state(self._h, np.ascontiguousarray(U, np.float64).ravel()))

    def set_state_device(self, ptr):
        check(lib().cfd2d_set_state_device(self._h, C.c_void_p(ptr)))

    def get_state_device(self, ptr):
        check(lib().cfd2d_get_state_device(self._h, C.c_void_p(ptr)))

    def state(self):
        U = np.zeros(4 * self.nc)
        check(lib().cfd2d_get_state(self._h, U))
        return U.reshape(4, self.nc)

    def flags(self):
        f = np.zeros(self.nc, np.uint32)
        check(lib().cfd2d_get_flags(self._h, f))
        return f

    def set_flags(self, f):
        check(lib().cfd2d_set_flags(self._h, np.ascontiguousarray(f, np.uint32)))

    def calc_tau(self):
        check(lib().cfd2d_calc_tau(self._h))

    def tau(self):
        t = C.c_double()
        ct = np.zeros(self.nc)
        check(lib().cfd2d_get_tau(self._h, C.byref(t), ct.ctypes.data_as(C.c_void_p)))
        return t.value, ct

    def set_tau(self, tau):
        check(lib().cfd2d_set_tau(self._h, tau))

    def set_bc_par(self, par):
        """Parameters of every non-periodic boundary object, (n_bc, 4)."""
        par = np.ascontiguousarray(par, np.float64).reshape(-1)
        if par.size != 4 * self.n_bc:
            raise ValueError("need %d x 4 boundary parameters" % self.n_bc)
        check(lib().cfd2d_set_bc_par(self._h, par if par.size else np.zeros(4)))

    def gradients(self):
        g = np.zeros(8 * self.nc)
        check(lib().cfd2d_gradients(self._h, g))
        return g.reshape(8, self.nc)

    def edge_fluxes(self):
        f = np.zeros(4 * self.ne)
        check(lib().cfd2d_edge_fluxes(self._h, f))
        return f.reshape(self.ne, 4)

    def residual(self):
        r = np.zeros(4 * self.nc)
        check(lib().cfd2d_residual(self._h, r))
        return r.reshape(4, self.nc)

    def step(self, n=1):
        check(lib().cfd2d_step(self._h, int(n)))

    def step_graph(self, n=1):
        check(lib().cfd2d_step_graph(self._h, int(n)))

    def sync(self):
        check(lib().cfd2d_sync(self._h))

    def status(self):
        hits, nonfinite = C.c_longlong(), C.c_int()
        check(lib().cfd2d_status(self._h, C.byref(hits), C.byref(nonfinite)))
        return hits.value, nonfinite.value

    def wall_force(self, edge_mask):
        fx, fy = C.c_double(), C.c_double()
        check(lib().cfd2d_wall_force(self._h, np.ascontiguousarray(edge_mask, np.int32), C.byref(fx), C.byref(fy)))
        return fx.value, fy.value

    def adjoint(self, nsteps, edge_mask, ckpt_every=20):
        """Runs nsteps steps and the reverse sweep for J = sum over masked
        edges of p(c1) * l * n_y; returns (J, dJ/dU0 as (4, nc), dJ/d(boundary
        parameters) as (n_bc, 4)).  The context ends at the final state.
        ckpt_every <= 0 lets the library choose the interval (all stages
        recorded when they fit in half of the free device memory)."""
        J = C.c_double()
        gU = np.zeros(4 * self.nc)
        gb = np.zeros(4 * max(self.n_bc, 1))
        check(lib().cfd2d_adjoint(self._h, int(nsteps), int(ckpt_every), np.ascontiguousarray(edge_mask, np.int32),
                                  C.byref(J), gU.ctypes.data_as(C.c_void_p), gb.ctypes.data_as(C.c_void_p)))
        return J.value, gU.reshape(4, self.nc), gb.reshape(-1, 4)[:self.n_bc]

    def adjoint_geometry(self, nsteps, edge_mask, ckpt_every=20):
        """adjoint() plus dJ/d(node coordinates) of the context's mesh as
        (n_nodes, 2); returns (J, dJ/dU0, dJ/dpar, dJ/dxy)."""
        J = C.c_double()
        gU = np.zeros(4 * self.nc)
        gb = np.zeros(4 * max(self.n_bc, 1))
        gx = np.zeros(2 * self.mesh.nn)
        check(lib().cfd2d_adjoint_geometry(self._h, self.mesh._h, int(nsteps), int(ckpt_every),
                                           np.ascontiguousarray(edge_mask, np.int32), C.byref(J),
                                           gU.ctypes.data_as(C.c_void_p), gb.ctypes.data_as(C.c_void_p), gx))
        return J.value, gU.reshape(4, self.nc), gb.reshape(-1, 4)[:self.n_bc], gx.reshape(-1, 2)

    def geometry_vjp_device(self, lsq, Ge, Gc, lny=None):
        """cfd2d_geometry_vjp_device: the mesh reverse of cfd2d_mesh_geometry_vjp
        on the GPU for Ge (13, n_edges), Gc (7, n_cells), lny (n_edges) or
        None; returns dxy (n_nodes, 2)."""
        Ge = np.ascontiguousarray(Ge, np.float64).ravel()
        Gc = np.ascontiguousarray(Gc, np.float64).ravel()
        ln = None if lny is None else np.ascontiguousarray(lny, np.float64)
        dxy = np.zeros(2 * self.mesh.nn)
        check(lib().cfd2d_geometry_vjp_device(self._h, self.mesh._h, int(bool(lsq)), Ge, Gc,
                                              None if ln is None else ln.ctypes.data_as(C.c_void_p), dxy))
        return dxy.reshape(-1, 2)

    def adjoint_graph(self, enable=True):
        """Record the adjoint's passes as graphs and replay them (default) or
        launch every kernel directly; the same results."""
        check(lib().cfd2d_adjoint_graph(self._h, 1 if enable else 0))

    def adjoint_objectives(self, n_cases):
        """Per-case objective of the last adjoint call (concatenated meshes)."""
        out = np.zeros(n_cases)
        check(lib().cfd2d_adjoint_objectives(self._h, out))
        return out

    def implicit_assemble(self, steady=False, cfl=0.5, tau=1.0e-3, outputs=True, out=None):
        """The implicit method's linear system for the current state as block
        rows: (blocks (nc, 4, 4, 4) with slot 0 the diagonal and slots 1..3 the
        faces in ascending edge order, cols (nc, 4) with -1 on boundary
        faces, rhs (nc, 4), ctau (nc,)).  outputs=False assembles on the
        device only (nothing copied back) and returns None; out = the four
        arrays to fill (e.g. from implicit_buffers: page-locked)."""
        nc = self.nc
        if not outputs:
            check(lib().cfd2d_implicit_assemble(self._h, int(bool(steady)), float(cfl), float(tau), None, None, None,
                                                None))
            return None
        if out is None:
            out = (np.zeros((nc, 4, 4, 4)), np.zeros((nc, 4), np.int32), np.zeros((nc, 4)), np.zeros(nc))
        blocks, cols, rhs, ctau = out
        for a, shape, dt in ((blocks, (nc, 4, 4, 4), np.float64), (cols, (nc, 4), np.int32), (rhs, (nc, 4), np.float64),
                             (ctau, (nc,), np.float64)):
            if a.shape != shape or a.dtype != dt or not a.flags.c_contiguous:
                raise ValueError("implicit_assemble: output must be C-contiguous %s %s" % (shape, np.dtype(dt).name))
        vp = C.c_void_p
        check(lib().cfd2d_implicit_assemble(self._h, int(bool(steady)), float(cfl), float(tau),
                                            blocks.ctypes.data_as(vp), cols.ctypes.data_as(vp),
                                            rhs.ctypes.data_as(vp), ctau.ctypes.data_as(vp)))
        return blocks, cols, rhs, ctau

    def implicit_buffers(self):
        """Page-locked output arrays for implicit_assemble(out=...)."""
        nc = self.nc
        return (pinned_array((nc, 4, 4, 4), np.float64), pinned_array((nc, 4), np.int32),
                pinned_array((nc, 4), np.float64), pinned_array((nc,), np.float64))

    def fuse_gradient(self, mode=True):
        """Where the gradients are formed: True/1 in the edge pass by each
        edge's lanes, 2 or "tiled" in the tiled edge pass (once per cell and
        tile, staged in LDS), False/0 in their own pass, None/-1 the default
        (edge pass unless a limiter is on); the same results."""
        if mode is None:
            m = -1
        elif mode == "tiled" or (not isinstance(mode, bool) and mode == 2):
            m = 2
        else:
            m = 1 if mode else 0
        check(lib().cfd2d_fuse_gradient(self._h, m))

    def stream_state(self, mode=True):
        """Whether the update pass reads and writes the conserved state
        non-temporally: True/1 yes, False/0 no, None/-1 the default (yes when
        the context's arrays exceed twice the 256 MB Infinity Cache); the same
        results."""
        m = -1 if mode is None else (1 if mode else 0)
        check(lib().cfd2d_stream_state(self._h, m))

    def autotune(self, trial_steps=3):
        """Times the step variants on the current state (restored afterwards)
        and keeps the fastest; returns dict(fuse, graph, stream, edge_shape,
        ms_per_step)."""
        f, g, st, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        check(lib().cfd2d_autotune(self._h, int(trial_steps), C.byref(f), C.byref(g), C.byref(st), C.byref(ms)))
        out = self.tuning()
        out.update(fuse=f.value, graph=g.value, stream=st.value, ms_per_step=ms.value)
        return out

    def edge_shape(self, shape):
        """Launch shape of the edge pass (0..3, cfd2d_edge_shape); same results."""
        check(lib().cfd2d_edge_shape(self._h, int(shape)))

    def tuning(self):
        f, g, st, sh = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().cfd2d_tuning(self._h, C.byref(f), C.byref(g), C.byref(st), C.byref(sh)))
        return dict(fuse=f.value, graph=g.value, stream=st.value, edge_shape=sh.value)

    def timing(self, enable=True, kernels=None):
        """Per-kernel HIP-event timing on or off (counters reset); kernels =
        names from KERNEL_NAMES to time (default all)."""
        check(lib().cfd2d_timing(self._h, 1 if enable else 0))
        mask = 0xFFFFFFFF if kernels is None else sum(1 << KERNEL_NAMES.index(k) for k in kernels)
        check(lib().cfd2d_timing_kernels(self._h, mask))

    def timing_read(self):
        out = {}
        for k, name in enumerate(KERNEL_NAMES):
            ms, n = C.c_double(), C.c_longlong()
            check(lib().cfd2d_timing_read(self._h, k, C.byref(ms), C.byref(n)))
            out[name] = (ms.value, n.value)
        return out

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value and _lib is not None:
            _lib.cfd2d_ctx_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()


class _Pinned:
    """Page-locked host allocation (cfd2d_host_alloc), freed with its last array."""

    def __init__(self, nbytes):
        self.ptr = C.c_void_p()
        check(lib().cfd2d_host_alloc(nbytes, C.byref(self.ptr)))
        self.__array_interface__ = {"shape": (nbytes,), "typestr": "|u1", "data": (self.ptr.value, False),
                                    "version": 3}

    def __del__(self):
        if self.ptr and _lib is not None:
            _lib.cfd2d_host_free(self.ptr)
            self.ptr = None


def pinned_array(shape, dtype):
    dt = np.dtype(dtype)
    n = int(np.prod(shape)) * dt.itemsize
    buf = _Pinned(max(n, 1))
    return np.asarray(buf)[:n].view(dt).reshape(shape)


class GpuSolver:
    """A cfd2d.case.Case on the GPU: mesh build, boundary mapping, initial
    state and time step as FVM_TVD::init does them (fvm_tvd.cpp:7-213)."""

    def __init__(self, case, calc_tau=True):
        self.case = case
        self.mesh = Mesh.from_trimesh(case.mesh, case.periodic_pairs())
        self.ctx = Context(self.mesh, case.config, case.group_bc(), case.cell_material())
        self.nc, self.ne = self.mesh.nc, self.mesh.ne
        self.ctx.set_state(case.initial_state())
        if calc_tau:
            self.ctx.calc_tau()

    def __getattr__(self, name):
        return getattr(self.ctx, name)

    def edge_mask(self, group):
        """int32 mask of the edges in the named edge group (objective edges)."""
        names = self.mesh.group_names(0)
        if group not in names:
            raise KeyError("no edge group %r (have %s)" % (group, names))
        eg = self.mesh.export()["edge_group"]
        return (eg == names.index(group)).astype(np.int32)

    def run(self, n, graph=False):
        if graph:
            self.ctx.step_graph(n)
        else:
            self.ctx.step(n)
        self.ctx.sync()

    def close(self):
        self.ctx.close()
        self.mesh.close()
