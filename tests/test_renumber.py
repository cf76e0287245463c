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
