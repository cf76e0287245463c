"""Seeded triangular meshes for the explicit finite-volume step.

Every generator returns a :class:`TriMesh`: node coordinates, counter-clockwise
triangles, boundary edges tagged with a named group, and a named group per
cell.  The node/triangle numbering is the one the UNV writer emits, so a mesh
passed as arrays and the same mesh read back from its UNV file give the same
cell and edge numbering.

Jitter comes from a counter-based generator (SplitMix64 of a per-stream key
plus the node index), so a mesh is a pure function of its sizes and seed and
can be rebuilt bit-for-bit anywhere.
"""
from __future__ import annotations

from dataclasses import dataclass, field
import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(x):
    """SplitMix64 finaliser (Steele, Lea & Flood 2014) on a uint64 array."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, stream: int, index) -> np.ndarray:
    """U[-1, 1) variates for the given indices of one (seed, stream) pair."""
    key = splitmix64(np.uint64((seed * 1000003 + stream) & 0xFFFFFFFFFFFFFFFF))
    with np.errstate(over="ignore"):
        z = splitmix64(key + np.asarray(index, dtype=np.uint64))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


@dataclass
class TriMesh:
    xy: np.ndarray                 # (V, 2) float64
    tri: np.ndarray                # (N, 3) int32, counter-clockwise
    bedges: np.ndarray             # (B, 2) int32 node pairs
    bedge_group: np.ndarray        # (B,) int32 -> edge_group_names
    edge_group_names: list
    cell_group: np.ndarray         # (N,) int32 -> cell_group_names
    cell_group_names: list
    periodic: list = field(default_factory=list)   # [(groupA, groupB, (dx, dy))]
    meta: dict = field(default_factory=dict)

    @property
    def n_nodes(self):
        return int(self.xy.shape[0])

    @property
    def n_cells(self):
        return int(self.tri.shape[0])

    @property
    def n_bedges(self):
        return int(self.bedges.shape[0])

    def centroids(self):
        p = self.xy[self.tri]
        return (p[:, 0, :] + p[:, 1, :] + p[:, 2, :]) / 3.0

    def group_index(self, name):
        return self.edge_group_names.index(name)


def _split_quads(nx, ny, node_id):
    """Two CCW triangles per quad, along the (i,j)-(i+1,j+1) diagonal, row-major."""
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    i = i.ravel()
    j = j.ravel()
    a = node_id(i, j)
    b = node_id(i + 1, j)
    c = node_id(i + 1, j + 1)
    d = node_id(i, j + 1)
    t = np.empty((nx * ny, 2, 3), dtype=np.int64)
    t[:, 0, :] = np.stack([a, b, c], axis=1)
    t[:, 1, :] = np.stack([a, c, d], axis=1)
    return t.reshape(-1, 3).astype(np.int32)


def _box_boundary(nx, ny, node_id):
    """Boundary node pairs of an (nx x ny) node-lattice box, grouped bottom/right/top/left."""
    i = np.arange(nx)
    j = np.arange(ny)
    bottom = np.stack([node_id(i, 0 * i), node_id(i + 1, 0 * i)], axis=1)
    right = np.stack([node_id(0 * j + nx, j), node_id(0 * j + nx, j + 1)], axis=1)
    top = np.stack([node_id(i + 1, 0 * i + ny), node_id(i, 0 * i + ny)], axis=1)
    left = np.stack([node_id(0 * j, j + 1), node_id(0 * j, j)], axis=1)
    return bottom, right, top, left


def rect_mesh(nx=200, ny=20, lx=1.0, ly=0.2, jitter=0.1, seed=1, split_x=0.5,
              cell_names=("cells_left", "cells_right")):
    """Jittered structured triangulation of [0,lx] x [0,ly].

    Interior nodes move by U(-jitter, jitter) times the cell size in each
    direction; boundary nodes slide along their side only; corners stay put.
    Cells whose centroid has x < split_x belong to cell_names[0], the rest to
    cell_names[1] (a Sod-type two-state initial condition).  Edge groups are
    edge_bottom, edge_right, edge_top, edge_left.
    """
    hx = lx / nx
    hy = ly / ny
    jj, ii = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    ii = ii.ravel()
    jj = jj.ravel()
    k = np.arange(ii.size)
    x = ii * hx
    y = jj * hy
    jx = uniform(seed, 0, k) * jitter * hx
    jy = uniform(seed, 1, k) * jitter * hy
    # nodes on the left/right sides keep x, nodes on the bottom/top keep y
    movex = (ii > 0) & (ii < nx)
    movey = (jj > 0) & (jj < ny)
    x = np.where(movex, x + jx, x)
    y = np.where(movey, y + jy, y)
    xy = np.stack([x, y], axis=1)

    def nid(i, j):
        return j * (nx + 1) + i

    tri = _split_quads(nx, ny, nid)
    b, r, t, l = _box_boundary(nx, ny, nid)
    bedges = np.concatenate([b, r, t, l]).astype(np.int32)
    bgroup = np.concatenate([np.full(len(b), 0), np.full(len(r), 1),
                             np.full(len(t), 2), np.full(len(l), 3)]).astype(np.int32)
    m = TriMesh(xy, tri, bedges, bgroup, ["edge_bottom", "edge_right", "edge_top", "edge_left"],
                np.zeros(tri.shape[0], np.int32), list(cell_names))
    cx = m.centroids()[:, 0]
    m.cell_group = np.where(cx < split_x, 0, 1).astype(np.int32)
    m.meta = dict(kind="rect", nx=nx, ny=ny, lx=lx, ly=ly, jitter=jitter, seed=seed)
    return m


def periodic_square(n=1024, jitter=0.1, seed=2, cell_name="fluid"):
    """Jittered triangulation of the unit square, periodic in x and y.

    Node (i, j) is jittered by the variate of (i mod n, j mod n), so the
    right/top copies of the seam nodes are exact translates of the left/bottom
    ones.  The edge groups edge_right/edge_left and edge_top/edge_bottom are
    paired by translation (1,0) and (0,1).
    """
    h = 1.0 / n
    jj, ii = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    ii = ii.ravel()
    jj = jj.ravel()
    k = (jj % n) * n + (ii % n)
    jx = uniform(seed, 0, k) * jitter * h
    jy = uniform(seed, 1, k) * jitter * h
    x = (ii % n) * h + jx + np.where(ii == n, 1.0, 0.0)
    y = (jj % n) * h + jy + np.where(jj == n, 1.0, 0.0)
    xy = np.stack([x, y], axis=1)

    def nid(i, j):
        return j * (n + 1) + i

    tri = _split_quads(n, n, nid)
    b, r, t, l = _box_boundary(n, n, nid)
    bedges = np.concatenate([b, r, t, l]).astype(np.int32)
    bgroup = np.concatenate([np.full(len(b), 0), np.full(len(r), 1),
                             np.full(len(t), 2), np.full(len(l), 3)]).astype(np.int32)
    m = TriMesh(xy, tri, bedges, bgroup, ["edge_bottom", "edge_right", "edge_top", "edge_left"],
                np.zeros(tri.shape[0], np.int32), [cell_name],
                periodic=[("edge_right", "edge_left", (1.0, 0.0)),
                          ("edge_top", "edge_bottom", (0.0, 1.0))])
    m.meta = dict(kind="periodic_square", n=n, jitter=jitter, seed=seed)
    return m


def bump_profile(x, height):
    """Circular-arc bump of the given height on 1 <= x <= 2 (zero elsewhere)."""
    x = np.asarray(x, dtype=np.float64)
    r = (0.25 + height * height) / (2.0 * height)
    inside = (x >= 1.0) & (x <= 2.0)
    arg = np.maximum(r * r - (x - 1.5) ** 2, 0.0)
    return np.where(inside, np.sqrt(arg) + (height - r), 0.0)


def bump_channel(ny=20, height=0.05, jitter=0.1, seed=3, cell_name="fluid"):
    """Channel [0,3] x [0,1] with a circular-arc bump on the bottom wall.

    An nx = 3*ny by ny structured grid is bent onto the bump and jittered
    (interior nodes in both directions, wall nodes along x with y put back on
    the wall, inlet/outlet nodes along y, corners fixed).  Edge groups:
    wall_bottom, bump, outlet, wall_top, inlet.
    """
    nx = 3 * ny
    jj, ii = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    ii = ii.ravel()
    jj = jj.ravel()
    k = np.arange(ii.size)
    hx = 3.0 / nx
    x = ii * hx
    jx = uniform(seed, 0, k) * jitter * hx
    jy = uniform(seed, 1, k) * jitter
    movex = (ii > 0) & (ii < nx)
    x = np.where(movex, x + jx, x)
    yb = bump_profile(x, height)
    eta = jj / ny
    hy = (1.0 - yb) / ny
    movey = (jj > 0) & (jj < ny)
    eta_y = yb + (1.0 - yb) * eta + np.where(movey, jy * hy, 0.0)
    xy = np.stack([x, eta_y], axis=1)

    def nid(i, j):
        return j * (nx + 1) + i

    tri = _split_quads(nx, ny, nid)
    b, r, t, l = _box_boundary(nx, ny, nid)
    bx = 0.5 * (xy[b[:, 0], 0] + xy[b[:, 1], 0])
    on_bump = (xy[b[:, 0], 0] >= 1.0) & (xy[b[:, 1], 0] <= 2.0) & (bx > 1.0) & (bx < 2.0)
    bgroup_b = np.where(on_bump, 1, 0)
    bedges = np.concatenate([b, r, t, l]).astype(np.int32)
    bgroup = np.concatenate([bgroup_b, np.full(len(r), 2), np.full(len(t), 3),
                             np.full(len(l), 4)]).astype(np.int32)
    m = TriMesh(xy, tri, bedges, bgroup, ["wall_bottom", "bump", "outlet", "wall_top", "inlet"],
                np.zeros(tri.shape[0], np.int32), [cell_name])
    m.meta = dict(kind="bump_channel", ny=ny, nx=nx, height=height, jitter=jitter, seed=seed)
    return m


def bump_batch(n_cases=256, min_cells=2000, max_cells=60000, seed=7, mach_range=(0.2, 3.0)):
    """Sizes, bump heights and inflow (Mach, angle) for a batch of bump-channel cases.

    Cell counts are log-uniform in [min_cells, max_cells]; each case's grid is
    the (3*ny x ny) channel with 6*ny^2 triangles closest to its draw; Mach
    uniform in mach_range, angle uniform in [-5, 5] degrees, bump height
    uniform in [0.02, 0.1].
    """
    idx = np.arange(n_cases)
    u0 = 0.5 * (uniform(seed, 10, idx) + 1.0)
    u1 = 0.5 * (uniform(seed, 11, idx) + 1.0)
    u2 = 0.5 * (uniform(seed, 12, idx) + 1.0)
    u3 = uniform(seed, 13, idx)
    cells = np.exp(math.log(min_cells) + u0 * (math.log(max_cells) - math.log(min_cells)))
    ny = np.maximum(np.rint(np.sqrt(cells / 6.0)).astype(int), 4)
    height = 0.02 + 0.08 * u1
    mach = mach_range[0] + (mach_range[1] - mach_range[0]) * u2
    angle = 5.0 * u3
    return [dict(ny=int(ny[c]), height=float(height[c]), mach=float(mach[c]),
                 angle_deg=float(angle[c]), seed=int(seed * 100003 + c)) for c in range(n_cases)]
