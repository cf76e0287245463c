"""The host half of the geometry sensitivities, cfd2d_mesh_geometry_vjp: the
reverse of the mesh builder (centroids, normals, lengths, Gauss points,
Heron areas, Green-Gauss weights n l and least-squares weights A^-1 d_k).

For random adjoint seeds of every geometric array the device reads, the
returned node-coordinate adjoint must be the gradient of the seeded sum
<seeds, geometry(xy)>, checked by central differences on the mesh rebuilt
with one node moved (h = 1e-6, 1e-6 relative; measured ~1e-8) -- every term of the reverse
is exercised (a dropped term, a wrong sign or a transposed index fails).
"""
import numpy as np
import pytest

from cfd2d import case as cases
from cfd2d import native


def geometry(tm, xy):
    """What the explicit kernels read, for node coordinates xy (the mesh
    builder's own numbers via cfd2d_mesh_create / export)."""
    m = native.Mesh.from_trimesh(tm.__class__(**{**tm.__dict__, "xy": xy}))
    ex = m.export()
    m.close()
    S, cx, cy = ex["cell_geom"][:, 0], ex["cell_geom"][:, 1], ex["cell_geom"][:, 2]
    nx, ny, l = ex["edge_geom"][:, 0], ex["edge_geom"][:, 1], ex["edge_geom"][:, 2]
    gp = ex["edge_geom"][:, 3:9]
    c1, c2 = ex["edge_int"][:, 2], ex["edge_int"][:, 3]
    sh = ex["edge_shift"]
    ce = ex["cell_edges"]
    ne, nc = len(nx), len(S)
    off = np.zeros((ne, 2, 2, 2))  # [edge][side][gauss point][x, y]
    c2s = np.where(c2 >= 0, c2, c1)
    for g in range(2):
        off[:, 0, g, 0] = gp[:, 2 + 2 * g] - cx[c1]
        off[:, 0, g, 1] = gp[:, 3 + 2 * g] - cy[c1]
        off[:, 1, g, 0] = np.where(c2 >= 0, gp[:, 2 + 2 * g] - (cx[c2s] + sh[:, 0]), 0.0)
        off[:, 1, g, 1] = np.where(c2 >= 0, gp[:, 3 + 2 * g] - (cy[c2s] + sh[:, 1]), 0.0)
    sg = np.where(c1[ce] == np.arange(nc)[:, None], 1.0, -1.0)
    gg = np.stack([sg * nx[ce] * l[ce], sg * ny[ce] * l[ce]], axis=2)  # [cell][face][x, y]
    # least squares: d_k to the neighbour (shifted) or twice the normal distance to the face
    d = np.zeros((nc, 3, 2))
    for k in range(3):
        e = ce[:, k]
        i = np.arange(nc)
        bnd = c2[e] < 0
        dn = (gp[e, 0] - cx) * nx[e] + (gp[e, 1] - cy) * ny[e]
        sgn = np.where(c1[e] == i, 1.0, -1.0)
        dxn = sgn * ((cx[c2s[e]] + sh[e, 0]) - cx[c1[e]])
        dyn = sgn * ((cy[c2s[e]] + sh[e, 1]) - cy[c1[e]])
        d[:, k, 0] = np.where(bnd, 2 * dn * nx[e], dxn)
        d[:, k, 1] = np.where(bnd, 2 * dn * ny[e], dyn)
    A = np.einsum("ikx,iky->ixy", d, d)
    w = np.einsum("ixy,iky->ikx", np.linalg.inv(A), d)
    return dict(nx=nx, ny=ny, l=l, off=off, S=S, gg=gg, w=w)


def seeded_sum(G, Ge, Gc, lny, lsq):
    ne = len(G["l"])
    tot = np.sum(Ge[0] * G["nx"] + Ge[1] * G["ny"] + Ge[2] * G["nx"] + Ge[3] * G["ny"])
    for s in range(2):
        for g in range(2):
            tot += np.sum(Ge[4 + 4 * s + 2 * g] * G["off"][:, s, g, 0] + Ge[5 + 4 * s + 2 * g] * G["off"][:, s, g, 1])
    tot += np.sum(Ge[12] * G["l"])
    tot += np.sum(Gc[0] * G["S"])
    W = G["w"] if lsq else G["gg"]
    for k in range(3):
        tot += np.sum(Gc[1 + 2 * k] * W[:, k, 0] + Gc[2 + 2 * k] * W[:, k, 1])
    tot += np.sum(lny * G["l"] * G["ny"])
    assert ne == Ge.shape[1]
    return tot


@pytest.mark.parametrize("lsq", [False, True], ids=["green_gauss", "lsq"])
def test_geometry_vjp_is_the_gradient(lsq):
    tm = cases.bump_channel(5, height=0.08, seed=4).mesh
    m = native.Mesh.from_trimesh(tm)
    ne, nc, nn = m.ne, m.nc, m.nn
    rng = np.random.default_rng(7)
    Ge = rng.standard_normal((13, ne))
    ex = m.export()
    Ge[8:12, ex["edge_int"][:, 3] < 0] = 0.0  # a boundary edge has no second side
    Gc = rng.standard_normal((7, nc))
    lny = np.where(rng.random(ne) < 0.2, rng.standard_normal(ne), 0.0)
    dxy = np.zeros(2 * nn)
    Ge, Gc = np.ascontiguousarray(Ge), np.ascontiguousarray(Gc)
    vp = lambda a: a.ctypes.data_as(native.C.c_void_p)
    native.check(native.lib().cfd2d_mesh_geometry_vjp(m._h, int(lsq), vp(Ge), vp(Gc), vp(lny), dxy))
    dxy = dxy.reshape(nn, 2)
    assert np.abs(dxy).max() > 1.0
    h = 1e-6
    nodes = list(range(0, nn, max(1, nn // 25))) + [nn - 1]
    worst = 0.0
    for v in nodes:
        for ax in range(2):
            xp = tm.xy.copy()
            xm = tm.xy.copy()
            xp[v, ax] += h
            xm[v, ax] -= h
            fd = (seeded_sum(geometry(tm, xp), Ge, Gc, lny, lsq) - seeded_sum(geometry(tm, xm), Ge, Gc, lny, lsq)) / (2 * h)
            err = abs(fd - dxy[v, ax]) / max(abs(fd), 1.0)
            worst = max(worst, err)
            assert err <= 1e-6, (v, ax, fd, dxy[v, ax])
    print("geometry vjp: worst relative error %.2e over %d coordinates" % (worst, 2 * len(nodes)))
    m.close()
