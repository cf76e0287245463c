"""Cell and node renumbering of a triangle mesh.

The mesh builder numbers edges in cell-discovery order, so the cell order
decides how far apart the cells and edges a kernel gathers together lie in
memory.  Meshes written by external generators (Salome, Triangle) come in
whatever order the generator used; `renumber(mesh, "hilbert")` sorts the
cells along a Hilbert curve through their centroids and numbers the nodes by
first use in that cell order, which keeps neighbours close (the role METIS
plays for the reference's partitioner, decomp.cpp:86-104, here for locality
on one device).  `"random"` shuffles (seeded), for measuring the effect.

A renumbered case is the same discrete problem: its results equal the
original's, mapped through the permutation, up to the order of floating-point
sums (each cell still adds its faces in ascending edge number, and the edge
numbers change).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import meshgen


def hilbert_index(ix, iy, bits=16):
    """Hilbert-curve position of integer points in [0, 2^bits)^2."""
    n = 1 << bits
    x = np.asarray(ix, dtype=np.int64).copy()
    y = np.asarray(iy, dtype=np.int64).copy()
    d = np.zeros_like(x)
    s = n >> 1
    while s > 0:
        rx = (x & s) > 0
        ry = (y & s) > 0
        d += s * s * ((3 * rx.astype(np.int64)) ^ ry.astype(np.int64))
        # rotate the quadrant so the curve stays continuous
        flip = (~ry) & rx
        x = np.where(flip, n - 1 - x, x)
        y = np.where(flip, n - 1 - y, y)
        swap = ~ry
        x, y = np.where(swap, y, x), np.where(swap, x, y)
        s >>= 1
    return d


def cell_order(mesh: meshgen.TriMesh, order="hilbert", seed=0):
    """Permutation p with new cell k = old cell p[k]."""
    if order == "random":
        return np.random.default_rng(seed).permutation(mesh.n_cells)
    if order != "hilbert":
        raise ValueError("order must be 'hilbert' or 'random'")
    c = mesh.centroids()
    lo, hi = c.min(axis=0), c.max(axis=0)
    span = np.where(hi > lo, hi - lo, 1.0)
    q = np.minimum(((c - lo) / span * 65535.0).astype(np.int64), 65535)
    return np.argsort(hilbert_index(q[:, 0], q[:, 1]), kind="stable")


def renumber(mesh: meshgen.TriMesh, order="hilbert", seed=0):
    """(renumbered mesh, cell permutation p: new cell k = old cell p[k])."""
    p = cell_order(mesh, order, seed)
    tri = mesh.tri[p]
    # nodes by first use in the new cell order
    first = np.full(mesh.n_nodes, -1, np.int64)
    flat = tri.ravel()
    seen_order = flat[np.sort(np.unique(flat, return_index=True)[1])]
    first[seen_order] = np.arange(seen_order.size)
    unused = np.nonzero(first < 0)[0]
    first[unused] = seen_order.size + np.arange(unused.size)
    inv_nodes = np.empty(mesh.n_nodes, np.int64)
    inv_nodes[first] = np.arange(mesh.n_nodes)
    new = dataclasses.replace(
        mesh,
        xy=mesh.xy[inv_nodes].copy(),
        tri=first[tri].astype(np.int32),
        bedges=first[mesh.bedges].astype(np.int32),
        cell_group=mesh.cell_group[p].copy(),
        meta=dict(mesh.meta, renumbered=order))
    return new, p


def renumber_case(case, order="hilbert", seed=0):
    """The case on the renumbered mesh, starting from the permuted initial
    state; returns (case, p)."""
    m, p = renumber(case.mesh, order, seed)
    U0 = case.initial_state()[:, p]
    return dataclasses.replace(case, mesh=m, U0=U0, name=case.name + "_" + order), p
