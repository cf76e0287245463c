"""Mesh renumbering (cfd2d/renumber.py): the curve is a continuous
bijection, a renumbered case is the same discrete problem (oracle states map
through the permutation to round-off), and Hilbert order restores locality."""
import numpy as np
import pytest

import oracle
from cfd2d import case as cases
from cfd2d import renumber


def test_hilbert_curve_is_a_continuous_bijection():
    bits = 4
    ix, iy = np.meshgrid(np.arange(1 << bits), np.arange(1 << bits), indexing="ij")
    ix, iy = ix.ravel(), iy.ravel()
    d = renumber.hilbert_index(ix, iy, bits=bits)
    assert sorted(d.tolist()) == list(range(1 << (2 * bits)))
    o = np.argsort(d)
    steps = np.abs(np.diff(ix[o])) + np.abs(np.diff(iy[o]))
    assert np.all(steps == 1)


@pytest.mark.parametrize("make", [lambda: cases.shock_tube(20, 4, mild=True),
                                  lambda: cases.periodic_pulse(12, seed=2)])
def test_renumbered_case_is_the_same_problem(make):
    c = make()
    s0 = oracle.OracleSolver(c)
    s0.run(15)
    U = s0.state()
    scale = np.abs(U).max(axis=1, keepdims=True)
    r, p = renumber.renumber_case(c, "random", seed=5)
    h, ph = renumber.renumber_case(r, "hilbert")
    for cc, perm in ((r, p), (h, p[ph])):
        s = oracle.OracleSolver(cc)
        s.run(15)
        assert np.max(np.abs(s.state() - U[:, perm]) / scale) <= 1e-13


def test_hilbert_order_restores_locality():
    c = cases.periodic_pulse(64, seed=2)
    r, _ = renumber.renumber_case(c, "random", seed=1)
    h, _ = renumber.renumber_case(r, "hilbert")

    def span(cc):
        ei = oracle.OracleSolver(cc).geometry["edge_int"]
        inner = ei[:, 3] >= 0
        return float(np.mean(np.abs(ei[inner, 2] - ei[inner, 3])))

    s_rand, s_hil = span(r), span(h)
    assert s_hil * 20 < s_rand, (s_hil, s_rand)  # 30x here; seam edges of the torus stay long


def _geometry(m):
    ex = m.export()
    return {k: ex[k] for k in ("cell_geom", "cell_nodes", "cell_edges", "edge_geom", "edge_int", "edge_group",
                               "edge_shift")}


@pytest.mark.parametrize("make", [lambda: cases.shock_tube(24, 5, mild=True),
                                  lambda: cases.periodic_pulse(10, seed=3)], ids=["tube", "periodic"])
def test_host_renumbering_matches_the_tool(make):
    """cfd2d_mesh_renumber (C++ host, the drop-in's path) gives the same cell
    permutation as cfd2d/renumber.py and a mesh bit-identical to the tool's
    renumbered mesh built from arrays: nodes by first use, groups, cell
    groups and periodic pairings carried over."""
    from cfd2d import native
    c = make()
    shuffled, _ = renumber.renumber_case(c, "random", seed=9)
    tool, p_tool = renumber.renumber_case(shuffled, "hilbert")
    m = native.Mesh.from_trimesh(shuffled.mesh, shuffled.periodic_pairs())
    r, p = m.renumber("hilbert")
    assert np.array_equal(p, p_tool)
    ref = native.Mesh.from_trimesh(tool.mesh, tool.periodic_pairs())
    a, b = _geometry(r), _geometry(ref)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(r.cell_groups(), ref.cell_groups())
    assert r.group_names(0) == ref.group_names(0) and r.group_names(1) == ref.group_names(1)
    assert np.array_equal(r.nodes(), ref.nodes())
    # the identity order rebuilds the same mesh
    same, pid = m.renumber("none")
    assert np.array_equal(pid, np.arange(m.nc))
    g0, g1 = _geometry(m), _geometry(same)
    for k in g0:
        assert np.array_equal(g0[k], g1[k]), k


def test_host_shuffle_and_locality():
    """The seeded shuffle is a permutation that scatters neighbours; Hilbert
    order brings them back close."""
    from cfd2d import native
    c = cases.periodic_pulse(48, seed=2)
    m = native.Mesh.from_trimesh(c.mesh, c.periodic_pairs())
    s, ps = m.renumber("shuffle", seed=4)
    assert sorted(ps.tolist()) == list(range(m.nc)) and not np.array_equal(ps, np.arange(m.nc))
    h, _ = s.renumber("hilbert")

    def span(mm):
        ei = mm.export()["edge_int"]
        inner = ei[:, 3] >= 0
        return float(np.mean(np.abs(ei[inner, 2] - ei[inner, 3])))

    assert span(h) * 20 < span(s)


def test_host_renumbering_rejects_a_batch():
    from cfd2d import native
    parts = [native.Mesh.from_trimesh(cases.shock_tube(6, 2, mild=True).mesh) for _ in range(2)]
    b = native.Mesh.concat(parts)
    with pytest.raises(native.Cfd2dError):
        b.renumber("hilbert")
