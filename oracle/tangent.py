"""Vectorised numpy model of the explicit step with forward-mode tangents.

TEST INFRASTRUCTURE ONLY: the tangent-linear reference the GPU adjoint is
checked against (<dJ/dU0, v> + <dJ/dpar, w> = dJ[v, w]).  It covers the
configurations the adjoint supports: LAX, HLLC or exact Godunov flux, Green-Gauss or
least-squares gradients, Heun RK2 or SSP-RK3, inlet / outlet / slip-wall /
pressure-outlet boundaries, periodic pairing, no limiter, no viscous terms,
a fixed time step.  Every formula follows the C oracle (fvm_oracle.c), which
is itself pinned to the reference; tests/test_adjoint_cpu.py checks that the
value part of this model equals the C oracle to round-off before the
tangents are trusted.

A `Dual` carries a value array and one tangent array; the model is written
once and evaluated on plain arrays (value) or Duals (value + tangent).
"""
from __future__ import annotations

import numpy as np

GAS_R = 8.314472


class Dual:
    __slots__ = ("v", "d")
    __array_ufunc__ = None  # ndarray (op) Dual defers to Dual's reflected operator

    def __init__(self, v, d=None):
        self.v = np.asarray(v, dtype=np.float64)
        self.d = np.zeros_like(self.v) if d is None else np.asarray(d, dtype=np.float64)

    def __add__(self, o):
        if isinstance(o, Dual):
            return Dual(self.v + o.v, self.d + o.d)
        return Dual(self.v + o, self.d)

    __radd__ = __add__

    def __sub__(self, o):
        if isinstance(o, Dual):
            return Dual(self.v - o.v, self.d - o.d)
        return Dual(self.v - o, self.d)

    def __rsub__(self, o):
        return Dual(o - self.v, -self.d)

    def __neg__(self):
        return Dual(-self.v, -self.d)

    def __mul__(self, o):
        if isinstance(o, Dual):
            return Dual(self.v * o.v, self.d * o.v + self.v * o.d)
        return Dual(self.v * o, self.d * o)

    __rmul__ = __mul__

    def __truediv__(self, o):
        if isinstance(o, Dual):
            return Dual(self.v / o.v, (self.d * o.v - self.v * o.d) / (o.v * o.v))
        return Dual(self.v / o, self.d / o)

    def __rtruediv__(self, o):
        return Dual(o / self.v, -o * self.d / (self.v * self.v))

    def __getitem__(self, k):
        return Dual(self.v[k], self.d[k])

    def __setitem__(self, k, val):
        if isinstance(val, Dual):
            self.v[k] = val.v
            self.d[k] = val.d
        else:
            self.v[k] = val
            self.d[k] = 0.0

    @property
    def shape(self):
        return self.v.shape


def val(x):
    return x.v if isinstance(x, Dual) else np.asarray(x)


def sqrt(x):
    if isinstance(x, Dual):
        s = np.sqrt(x.v)
        return Dual(s, x.d / (2.0 * s))
    return np.sqrt(x)


def expd(x):
    if isinstance(x, Dual):
        e = np.exp(x.v)
        return Dual(e, e * x.d)
    return np.exp(x)


def logd(x):
    if isinstance(x, Dual):
        return Dual(np.log(x.v), x.d / x.v)
    return np.log(x)


def absd(x):
    if isinstance(x, Dual):
        sg = np.where(x.v < 0, -1.0, 1.0)
        return Dual(np.abs(x.v), sg * x.d)
    return np.abs(x)


def where(c, a, b):
    if isinstance(a, Dual) or isinstance(b, Dual):
        a = a if isinstance(a, Dual) else Dual(np.broadcast_to(a, c.shape).copy())
        b = b if isinstance(b, Dual) else Dual(np.broadcast_to(b, c.shape).copy())
        return Dual(np.where(c, a.v, b.v), np.where(c, a.d, b.d))
    return np.where(c, a, b)


def stack(xs):
    if any(isinstance(x, Dual) for x in xs):
        xs = [x if isinstance(x, Dual) else Dual(x) for x in xs]
        return Dual(np.stack([x.v for x in xs]), np.stack([x.d for x in xs]))
    return np.stack(xs)


def zeros_like(x, shape):
    if isinstance(x, Dual):
        return Dual(np.zeros(shape))
    return np.zeros(shape)


class Model:
    """The explicit step on an exported oracle mesh (OracleSolver.geometry)."""

    def __init__(self, solver):
        case = solver.case
        c = case.config
        g = solver.geometry
        if c.flux not in ("LAX", "HLLC", "GODUNOV") or c.limiter != "NONE" or c.viscous:
            raise ValueError("tangent model: LAX/HLLC/GODUNOV, no limiter, inviscid only")
        self.flux, self.lsq, self.rk = c.flux, c.gradient in ("LSQ", "LEAST_SQUARES"), c.rk_stages
        self.gam_flux = c.gam_flux
        mat = c.materials[0]
        self.M, self.Cv = mat.M, mat.Cp - GAS_R / mat.M
        self.gam = mat.Cp / self.Cv
        self.nc, self.ne = solver.nc, solver.ne
        self.S = g["cell_geom"][:, 0].copy()
        self.cx, self.cy = g["cell_geom"][:, 1].copy(), g["cell_geom"][:, 2].copy()
        self.ce = g["cell_edges"].copy()
        self.c1, self.c2 = g["edge_int"][:, 2].copy(), g["edge_int"][:, 3].copy()
        self.nx, self.ny, self.l = g["edge_geom"][:, 0].copy(), g["edge_geom"][:, 1].copy(), g["edge_geom"][:, 2].copy()
        self.gp = g["edge_geom"][:, 3:9].copy()
        self.sh = g["edge_shift"].copy()
        self._cell_nodes = g["cell_nodes"].copy()
        self._n1, self._n2 = g["edge_int"][:, 0].copy(), g["edge_int"][:, 1].copy()
        self.nx0, self.ny0 = self.nx.copy(), self.ny.copy()
        self.ebc = solver.edge_bc.copy()
        bcs = [b for b in c.boundaries if b.type != "BOUND_PERIODIC"]
        self.bctype = np.array([b.code for b in bcs] or [0])
        self.bcpar0 = np.array([list(b.params()) for b in bcs] or [[0, 0, 0, 0]], dtype=np.float64)
        self.inner = self.c2 >= 0
        self.bnd = ~self.inner
        cc2 = np.where(self.inner, self.c2, self.c1)
        # Gauss-point offsets from each side (periodic shift on c2)
        self.d1 = [np.stack([self.gp[:, 2 + 2 * k] - self.cx[self.c1], self.gp[:, 3 + 2 * k] - self.cy[self.c1]]) for k in range(2)]
        self.d2 = [np.stack([self.gp[:, 2 + 2 * k] - (self.cx[cc2] + self.sh[:, 0]),
                             self.gp[:, 3 + 2 * k] - (self.cy[cc2] + self.sh[:, 1])]) for k in range(2)]
        self.side = (self.c1[self.ce] != np.arange(self.nc)[:, None])  # True where the cell is c2
        self.nbr = np.where(self.side, self.c1[self.ce], self.c2[self.ce])  # -1 on boundary faces
        if self.lsq:
            self._lsq_weights()
        self.dtS = None

    def set_geometry(self, xy, dxy):
        """Geometry rebuilt from node coordinates xy (n_nodes, 2) carrying the
        tangent dxy, with the expressions of the mesh builder (centroid
        ((x0 + x1) + x2) / 3; normal (y2 - y1, x1 - x2) / l, oriented as
        exported; Gauss points mid -/+ (x2 - x1) / (2 sqrt 3); Heron area over
        the cell's edges in order; least-squares weights A^-1 d_k): every
        geometric array becomes a Dual, so a run gives dJ[dxy]."""
        X = Dual(np.asarray(xy, dtype=np.float64)[:, 0].copy(), np.asarray(dxy, dtype=np.float64)[:, 0].copy())
        Y = Dual(np.asarray(xy, dtype=np.float64)[:, 1].copy(), np.asarray(dxy, dtype=np.float64)[:, 1].copy())
        cn = self._cell_nodes
        self.cx = (X[cn[:, 0]] + X[cn[:, 1]] + X[cn[:, 2]]) / 3.0
        self.cy = (Y[cn[:, 0]] + Y[cn[:, 1]] + Y[cn[:, 2]]) / 3.0
        a, b = self._n1, self._n2
        x1, y1, x2, y2 = X[a], Y[a], X[b], Y[b]
        rx, ry = y2 - y1, x1 - x2
        l = sqrt(rx * rx + ry * ry)
        sg = np.where(val(rx) * self.nx0 + val(ry) * self.ny0 >= 0.0, 1.0, -1.0)
        self.nx, self.ny, self.l = (rx / l) * sg, (ry / l) * sg, l
        sq3 = 1.0 / np.sqrt(3.0)
        mx, my = (x1 + x2) / 2.0, (y1 + y2) / 2.0
        cols = [mx, my, mx - sq3 * (x2 - x1) / 2.0, my - sq3 * (y2 - y1) / 2.0,
                mx + sq3 * (x2 - x1) / 2.0, my + sq3 * (y2 - y1) / 2.0]
        self.gp = Dual(np.stack([c.v for c in cols], axis=1), np.stack([c.d for c in cols], axis=1))
        ea, eb, ec = (self.l[self.ce[:, k]] for k in range(3))
        pp = (ea + eb + ec) / 2.0
        self.S = sqrt(pp * (pp - ea) * (pp - eb) * (pp - ec))
        self._offsets()
        if self.lsq:
            self._lsq_weights_dual()

    def _offsets(self):
        cc2 = np.where(self.inner, self.c2, self.c1)
        gp = self.gp
        self.d1 = [stack([gp[:, 2 + 2 * k] - self.cx[self.c1], gp[:, 3 + 2 * k] - self.cy[self.c1]]) for k in range(2)]
        self.d2 = [stack([gp[:, 2 + 2 * k] - (self.cx[cc2] + self.sh[:, 0]),
                          gp[:, 3 + 2 * k] - (self.cy[cc2] + self.sh[:, 1])]) for k in range(2)]

    def _lsq_weights_dual(self):
        nc = self.nc
        me = np.arange(nc)
        dx, dy = [], []
        for k in range(3):
            e = self.ce[:, k]
            c1, c2 = self.c1[e], self.c2[e]
            bnd = c2 < 0
            mx, my = self.gp[e, 0], self.gp[e, 1]
            nx, ny = self.nx[e], self.ny[e]
            dn = (mx - self.cx) * nx + (my - self.cy) * ny
            c2s = np.where(bnd, me, c2)
            ax = (self.cx[c2s] + self.sh[e, 0]) - self.cx[c1]
            ay = (self.cy[c2s] + self.sh[e, 1]) - self.cy[c1]
            bx = self.cx[c1] - (self.cx[c2s] + self.sh[e, 0])
            by = self.cy[c1] - (self.cy[c2s] + self.sh[e, 1])
            is1 = me == c1
            dx.append(where(bnd, 2.0 * dn * nx, where(is1, ax, bx)))
            dy.append(where(bnd, 2.0 * dn * ny, where(is1, ay, by)))
        a11 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]
        a12 = dx[0] * dy[0] + dx[1] * dy[1] + dx[2] * dy[2]
        a22 = dy[0] * dy[0] + dy[1] * dy[1] + dy[2] * dy[2]
        det = a11 * a22 - a12 * a12
        wx = [(a22 * dx[k] - a12 * dy[k]) / det for k in range(3)]
        wy = [(a11 * dy[k] - a12 * dx[k]) / det for k in range(3)]
        v = np.stack([np.stack([wx[k].v, wy[k].v], axis=1) for k in range(3)], axis=1)
        d = np.stack([np.stack([wx[k].d, wy[k].d], axis=1) for k in range(3)], axis=1)
        self.w = Dual(v, d)

    def _lsq_weights(self):
        nc = self.nc
        d = np.zeros((nc, 3, 2))
        for k in range(3):
            e = self.ce[:, k]
            c1, c2 = self.c1[e], self.c2[e]
            me = np.arange(nc)
            bnd = c2 < 0
            mx, my = self.gp[e, 0], self.gp[e, 1]
            dn = (mx - self.cx) * self.nx[e] + (my - self.cy) * self.ny[e]
            c2s = np.where(bnd, me, c2)
            ax = (self.cx[c2s] + self.sh[e, 0]) - self.cx[c1]
            ay = (self.cy[c2s] + self.sh[e, 1]) - self.cy[c1]
            bx = self.cx[c1] - (self.cx[c2s] + self.sh[e, 0])
            by = self.cy[c1] - (self.cy[c2s] + self.sh[e, 1])
            is1 = me == c1
            d[:, k, 0] = np.where(bnd, 2.0 * dn * self.nx[e], np.where(is1, ax, bx))
            d[:, k, 1] = np.where(bnd, 2.0 * dn * self.ny[e], np.where(is1, ay, by))
        a11 = (d[:, :, 0] ** 2).sum(1)
        a12 = (d[:, :, 0] * d[:, :, 1]).sum(1)
        a22 = (d[:, :, 1] ** 2).sum(1)
        det = a11 * a22 - a12 * a12
        self.w = np.stack([(a22[:, None] * d[:, :, 0] - a12[:, None] * d[:, :, 1]) / det[:, None],
                           (a11[:, None] * d[:, :, 1] - a12[:, None] * d[:, :, 0]) / det[:, None]], axis=2)

    # -- pieces -----------------------------------------------------------
    def prim(self, U):
        r = U[0]
        u = U[1] / r
        v = U[2] / r
        E = U[3] / r
        e = E - 0.5 * (u * u + v * v)
        p = r * e * (self.gam - 1)
        return [r, p, u, v]

    def ghost(self, e, L, C, par):
        """Ghost of state L (list r,p,u,v over boundary edges e) with cell-centre state C."""
        t = self.bctype[self.ebc[e]]
        nx, ny = self.nx[e], self.ny[e]
        Tc = (C[1] / (C[0] * (self.gam - 1))) / self.Cv
        czc = sqrt(self.gam * C[1] / C[0])
        pr = [par[k][self.ebc[e]] if isinstance(par[k], Dual) else par[k][self.ebc[e]] for k in range(4)]
        Un = L[2] * nx + L[3] * ny
        uw = L[2] - nx * Un * 2.0
        vw = L[3] - ny * Un * 2.0
        inlet, outlet, wall, pout = t == 1, t == 2, t == 3, t == 4
        u = where(inlet, pr[0], where(wall, uw, L[2]))
        v = where(inlet, pr[1], where(wall, vw, L[3]))
        T = where(inlet, pr[2], Tc)
        un = L[2] * nx + L[3] * ny
        sub = absd(val(un)) < val(czc)
        p = where(inlet, pr[3], where(pout & sub, pr[0], L[1]))
        r = where(pout, L[0], p * self.M / (T * GAS_R))
        return [r, p, u, v]

    def gradients(self, P, par):
        nc = self.nc
        G = []
        i = np.arange(nc)
        nb = [[] for _ in range(3)]
        for k in range(3):
            e = self.ce[:, k]
            j = self.nbr[:, k]
            bnd = j < 0
            js = np.where(bnd, i, j)
            ent = [P[q][js] for q in range(4)]
            if bnd.any():
                eb = e[bnd]
                cen = [P[q][bnd] for q in range(4)]
                gh = self.ghost(eb, cen, cen, par)
                for q in range(4):
                    ent[q] = _put(ent[q], bnd, gh[q])
            nb[k] = ent
        out = []
        for q in range(4):
            gx = 0.0
            gy = 0.0
            for k in range(3):
                if self.lsq:
                    dphi = nb[k][q] - P[q]
                    gx = gx + self.w[:, k, 0] * dphi
                    gy = gy + self.w[:, k, 1] * dphi
                else:
                    e = self.ce[:, k]
                    s = (P[q] + nb[k][q]) / 2
                    tx = s * self.nx[e] * self.l[e]
                    ty = s * self.ny[e] * self.l[e]
                    sg = np.where(self.side[:, k], -1.0, 1.0)
                    gx = gx + sg * tx
                    gy = gy + sg * ty
            if not self.lsq:
                gx = gx / self.S
                gy = gy / self.S
            out.append((gx, gy))
        return out

    def godunov(self, L, R, nx, ny):
        """FVM_TVD::calcFlux (src/methods/fvm_tvd.cpp:602-622) with rim_orig
        (src/global.cpp:232-405) as the C oracle restates it, row-vectorised:
        every row iterates Newton until its own stop test holds (rows that
        stopped keep their P), branches per row on the values, powers as
        exp(log(x) k)."""
        G = self.gam_flux
        AGAM = G - 1.0
        DGAM, GGAM, HGAM, FGAM = 2.0 / AGAM, np.sqrt(G * AGAM), AGAM / 2.0, 3.0 * G - 1.0
        OGAM, QGAM = AGAM / (2.0 * G), G + 1.0
        PGAM, RGAM, SGAM, TGAM = QGAM / (2.0 * G), 4.0 * G, G * AGAM, QGAM / 2.0
        eps = 1.0e-5
        unl = L[2] * nx + L[3] * ny
        unr = R[2] * nx + R[3] * ny
        utl = L[2] * ny - L[3] * nx
        utr = R[2] * ny - R[3] * nx
        RB, PB, UB, VB = L[0], L[1], unl, utl
        RE, PE, UE, VE = R[0], R[1], unr, utr
        CB = sqrt(G * PB / RB)
        CE = sqrt(G * PE / RE)
        EB = CB * CB / SGAM
        EE = CE * CE / SGAM
        RCB, RCE = RB * CB, RE * CE
        DU = UB - UE
        vac = val(DU) < val(-2.0 * (CB + CE) / AGAM)
        P = (PB * RCE + PE * RCB + DU * RCB * RCE) / (RCB + RCE)
        active = ~vac
        it = 0
        while active.any():
            P = where(active & (val(P) < eps), eps, P)
            PPB, PPE = P / PB, P / PE
            rb, re_ = val(PB) > val(P), val(PE) > val(P)
            ZFB = CB * expd(logd(where(rb, PPB, 1.0)) * OGAM)
            PKB = PGAM * PPB + OGAM
            ZNB = RCB * sqrt(where(rb, 1.0, PKB))
            F1 = where(rb, DGAM * (ZFB - CB), (P - PB) / ZNB)
            FS1 = where(rb, ZFB / (G * P), (QGAM * PPB + FGAM) / (RGAM * ZNB * PKB))
            ZFE = CE * expd(logd(where(re_, PPE, 1.0)) * OGAM)
            PKE = PGAM * PPE + OGAM
            ZNE = RCE * sqrt(where(re_, 1.0, PKE))
            F2 = where(re_, DGAM * (ZFE - CE), (P - PE) / ZNE)
            FS2 = where(re_, ZFE / (G * P), (QGAM * PPE + FGAM) / (RGAM * ZNE * PKE))
            res = DU - F1 - F2
            P = where(active, P + res / (FS1 + FS2), P)
            active = active & (np.abs(val(res)) > eps)
            it += 1
            if it > 500:
                raise RuntimeError("tangent model: rim_orig did not converge in 500 iterations")
        # the star states (the vacuum rows take the vacuum branch below)
        Ps = where(vac, 1.0, P)
        rb, re_ = val(PB) > val(Ps), val(PE) > val(Ps)
        ZFB = CB * expd(logd(where(rb, Ps / PB, 1.0)) * OGAM)
        ZFE = CE * expd(logd(where(re_, Ps / PE, 1.0)) * OGAM)
        EFr = ZFB * ZFB / SGAM
        UFr = UB + DGAM * (CB - ZFB)
        RFr = Ps / (AGAM * EFr)
        Db = UB - sqrt(where(rb, 1.0, (TGAM * Ps + HGAM * PB) / RB))
        RUBD = RB * (UB - Db)
        RFs = RUBD * RUBD / (PB - Ps + RUBD * (UB - Db))
        UFs = Db + RUBD / RFs
        EFs = Ps / (AGAM * RFs)
        ESr = ZFE * ZFE / SGAM
        USr = UE - DGAM * (CE - ZFE)
        RSr = Ps / (AGAM * ESr)
        De = UE + sqrt(where(re_, 1.0, (TGAM * Ps + HGAM * PE) / RE))
        RUED = RE * (UE - De)
        RSs = RUED * RUED / (PE - Ps + RUED * (UE - De))
        USs = De + RUED / RSs
        ESs = Ps / (AGAM * RSs)
        zero = where(vac, 0.0, 0.0 * Ps)
        RF = where(vac, zero, where(rb, RFr, RFs))
        EF = where(vac, zero, where(rb, EFr, EFs))
        UF = where(vac, zero, where(rb, UFr, UFs))
        SBL = where(vac, UB - CB, where(rb, UB - CB, Db))
        SFL = where(vac, UB + 2.0 * CB / AGAM, where(rb, UFr - ZFB, Db))
        RS = where(vac, zero, where(re_, RSr, RSs))
        ES = where(vac, zero, where(re_, ESr, ESs))
        US = where(vac, zero, where(re_, USr, USs))
        SEL = where(vac, UE + CE, where(re_, UE + CE, De))
        SSL = where(vac, UE - 2.0 * CE / AGAM, where(re_, USr + ZFE, De))
        # sampling at x/t = 0
        sel, sbl, ssl, sfl, us = val(SEL) <= 0.0, val(SBL) >= 0.0, val(SSL) >= 0.0, val(SFL) <= 0.0, val(US) >= 0.0
        UIl = (UB + DGAM * GGAM * sqrt(EB)) / (1 + DGAM)
        EIl = (UIl * UIl) / SGAM
        RIl = RB * expd(logd(EIl / EB) * (1 / AGAM))
        UIr = (UE - DGAM * GGAM * sqrt(EE)) / (1 + DGAM)
        EIr = (UIr * UIr) / SGAM
        RIr = RE * expd(logd(EIr / EE) * (1 / AGAM))
        fanl = val(SFL) > 0.0
        star = ssl & sfl
        pick = lambda e, b, f, s, fl, fr: where(sel, e, where(sbl, b, where(star, where(us, f, s), where(fanl, fl, fr))))
        RI = pick(RE, RB, RF, RS, RIl, RIr)
        EI = pick(EE, EB, EF, ES, EIl, EIr)
        UI = pick(UE, UB, UF, US, UIl, UIr)
        VI = pick(VE, VB, VB, VE, VB, VE)
        PI = AGAM * EI * RI
        UN, UT = UI, VI
        ux = UN * nx + UT * ny
        vy = UN * ny - UT * nx
        f0 = RI * UN
        return [f0, f0 * ux + PI * nx, f0 * vy + PI * ny, (RI * (EI + 0.5 * (ux * ux + vy * vy)) + PI) * UN]

    def numflux(self, L, R, nx, ny):
        G = self.gam_flux
        if self.flux == "GODUNOV":
            return self.godunov(L, R, nx, ny)
        unl = L[2] * nx + L[3] * ny
        unr = R[2] * nx + R[3] * ny
        if self.flux == "LAX":
            al = absd(unl) + sqrt(G * L[1] / L[0])
            ar = absd(unr) + sqrt(G * R[1] / R[0])
            alpha = where(val(al) > val(ar), al, ar)
            rol, rul, rvl = L[0], L[0] * L[2], L[0] * L[3]
            rel = L[1] / (G - 1.0) + 0.5 * L[0] * (L[2] * L[2] + L[3] * L[3])
            ror, rur, rvr = R[0], R[0] * R[2], R[0] * R[3]
            rer = R[1] / (G - 1.0) + 0.5 * R[0] * (R[2] * R[2] + R[3] * R[3])
            frl = rol * unl
            frr = ror * unr
            return [0.5 * (frr + frl - alpha * (ror - rol)),
                    0.5 * (frr * R[2] + frl * L[2] + (R[1] + L[1]) * nx - alpha * (rur - rul)),
                    0.5 * (frr * R[3] + frl * L[3] + (R[1] + L[1]) * ny - alpha * (rvr - rvl)),
                    0.5 * ((rer + R[1]) * unr + (rel + L[1]) * unl - alpha * (rer - rel))]
        # HLLC (Toro 10.4-10.6), Davis wave speeds
        cl = sqrt(G * L[1] / L[0])
        cr = sqrt(G * R[1] / R[0])
        El = L[1] / (G - 1.0) + 0.5 * L[0] * (L[2] * L[2] + L[3] * L[3])
        Er = R[1] / (G - 1.0) + 0.5 * R[0] * (R[2] * R[2] + R[3] * R[3])
        a1, a2 = unl - cl, unr - cr
        sl = where(val(a1) < val(a2), a1, a2)
        b1, b2 = unl + cl, unr + cr
        sr = where(val(b1) > val(b2), b1, b2)
        fl = [L[0] * unl, L[0] * L[2] * unl + L[1] * nx, L[0] * L[3] * unl + L[1] * ny, (El + L[1]) * unl]
        fr = [R[0] * unr, R[0] * R[2] * unr + R[1] * nx, R[0] * R[3] * unr + R[1] * ny, (Er + R[1]) * unr]
        ml = L[0] * (sl - unl)
        mr = R[0] * (sr - unr)
        ss = (R[1] - L[1] + L[0] * unl * (sl - unl) - R[0] * unr * (sr - unr)) / (ml - mr)
        kl = ml / (sl - ss)
        kr = mr / (sr - ss)
        usl = [kl, kl * (L[2] + (ss - unl) * nx), kl * (L[3] + (ss - unl) * ny),
               kl * (El / L[0] + (ss - unl) * (ss + L[1] / (L[0] * (sl - unl))))]
        usr = [kr, kr * (R[2] + (ss - unr) * nx), kr * (R[3] + (ss - unr) * ny),
               kr * (Er / R[0] + (ss - unr) * (ss + R[1] / (R[0] * (sr - unr))))]
        ul = [L[0], L[0] * L[2], L[0] * L[3], El]
        ur = [R[0], R[0] * R[2], R[0] * R[3], Er]
        out = []
        for q in range(4):
            fsl = fl[q] + sl * (usl[q] - ul[q])
            fsr = fr[q] + sr * (usr[q] - ur[q])
            out.append(where(val(sl) >= 0, fl[q], where(val(sr) <= 0, fr[q],
                                                      where(val(ss) >= 0, fsl, fsr))))
        return out

    def residual(self, U, par):
        P = self.prim(U)
        G = self.gradients(P, par)
        ne = self.ne
        c1 = self.c1
        c2s = np.where(self.inner, self.c2, self.c1)
        acc = None
        for k in range(2):
            d1, d2 = self.d1[k], self.d2[k]
            L = [P[q][c1] + (G[q][0][c1] * d1[0] + G[q][1][c1] * d1[1]) for q in range(4)]
            R = [P[q][c2s] + (G[q][0][c2s] * d2[0] + G[q][1][c2s] * d2[1]) for q in range(4)]
            if self.bnd.any():
                eb = np.nonzero(self.bnd)[0]
                Lb = [L[q][self.bnd] for q in range(4)]
                Cb = [P[q][c1[self.bnd]] for q in range(4)]
                gh = self.ghost(eb, Lb, Cb, par)
                for q in range(4):
                    R[q] = _put(R[q], self.bnd, gh[q])
            f = self.numflux(L, R, self.nx, self.ny)
            acc = f if acc is None else [acc[q] + f[q] for q in range(4)]
        hl = self.l * 0.5
        F = [acc[q] * hl for q in range(4)]
        Rs = []
        for q in range(4):
            r = 0.0
            for k in range(3):
                e = self.ce[:, k]
                sg = np.where(self.side[:, k], 1.0, -1.0)
                r = r + sg * F[q][e]
            Rs.append(r)
        return Rs

    def step(self, U, par):
        U0 = U
        if self.rk == 3:
            R = self.residual(U0, par)
            U1 = [U0[q] + self.dtS * R[q] for q in range(4)]
            R = self.residual(U1, par)
            U2 = [0.75 * U0[q] + 0.25 * (U1[q] + self.dtS * R[q]) for q in range(4)]
            R = self.residual(U2, par)
            return [(1.0 / 3.0) * U0[q] + (2.0 / 3.0) * (U2[q] + self.dtS * R[q]) for q in range(4)]
        R = self.residual(U0, par)
        U1 = [U0[q] + self.dtS * R[q] for q in range(4)]
        R = self.residual(U1, par)
        return [0.5 * (U0[q] + (U1[q] + self.dtS * R[q])) for q in range(4)]

    def objective(self, U, mask):
        P = self.prim(U)
        e = np.nonzero(mask)[0]
        p = P[1][self.c1[e]]
        t = p * self.l[e] * self.ny[e]
        s = 0.0
        for k in range(len(e)):
            s = s + t[k]
        return s

    def run(self, U0, par, nsteps, tau):
        self.dtS = tau / self.S
        U = [U0[q] for q in range(4)]
        for _ in range(nsteps):
            U = self.step(U, par)
        return U


def _put(arr, mask, vals):
    """arr with arr[mask] replaced by vals (arrays or Duals)."""
    if isinstance(arr, Dual) or isinstance(vals, Dual):
        a = arr if isinstance(arr, Dual) else Dual(arr.copy())
        a = Dual(a.v.copy(), a.d.copy())
        a[mask] = vals
        return a
    a = arr.copy()
    a[mask] = vals
    return a


def tangent(solver, U0, dU0, dpar, nsteps, tau, mask):
    """(J, dJ) for the state U0 (4, N) with tangent dU0 and boundary-parameter
    tangent dpar (n_bc, 4), after nsteps steps of fixed time step tau."""
    m = Model(solver)
    par = [Dual(m.bcpar0[:, k].copy(), np.asarray(dpar, dtype=np.float64)[:, k].copy()) for k in range(4)]
    U = [Dual(U0[q].copy(), dU0[q].copy()) for q in range(4)]
    U = m.run(U, par, nsteps, tau)
    J = m.objective(U, mask)
    return float(J.v), float(J.d)


def tangent_geometry(solver, xy, dxy, nsteps, tau, mask, U0=None):
    """(J, dJ) for node coordinates xy (n_nodes, 2) moving along dxy, the
    state U0 (default the case's) and boundary parameters held, after
    nsteps steps of fixed time step tau (the geometry rebuilt as Duals)."""
    m = Model(solver)
    m.set_geometry(xy, dxy)
    par = [m.bcpar0[:, k].copy() for k in range(4)]
    U0 = solver.case.initial_state() if U0 is None else U0
    U = m.run([Dual(U0[q].copy()) for q in range(4)], par, nsteps, tau)
    J = m.objective(U, mask)
    return float(J.v), float(J.d)


def forward(solver, U0, nsteps, tau):
    m = Model(solver)
    par = [m.bcpar0[:, k].copy() for k in range(4)]
    U = m.run([U0[q].copy() for q in range(4)], par, nsteps, tau)
    return np.stack(U)
