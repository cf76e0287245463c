"""Pins the C oracle to the reference: golden vectors from the compiled
reference (tests/golden/make_golden.py) plus textbook Riemann solutions."""
import hashlib
import math
import os
import tempfile

import numpy as np
import pytest

import golden_cases as gc
import oracle
from cfd2d import meshgen, unv


@pytest.mark.parametrize("name", gc.CASES)
def test_mesh_regenerates_same_unv(name):
    s = gc.spec(name)
    m = meshgen.rect_mesh(**s["mesh"])
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "mesh.unv")
        unv.write_unv(m, p)
        with open(p, "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == s["unv_sha256"]


@pytest.fixture(scope="module", params=gc.CASES)
def golden_run(request):
    name = request.param
    return name, gc.case(name), gc.load(name)


def test_geometry_bitwise(golden_run):
    name, case, g = golden_run
    s = oracle.OracleSolver(case)
    ex = s.geometry
    spec = gc.spec(name)
    for k in ("cell_geom", "cell_nodes", "cell_edges", "edge_geom", "edge_int"):
        if spec["full"]:
            assert np.array_equal(ex[k], g[k]), k
        else:
            assert hashlib.sha256(np.ascontiguousarray(ex[k]).tobytes()).hexdigest() == spec["geometry_sha256"][k], k


def test_first_stage_bitwise(golden_run):
    """Time step, U0, gradients, per-edge fluxes and the stage-1 residual."""
    name, case, g = golden_run
    s = oracle.OracleSolver(case)
    tau, ct = s.tau()
    assert tau == g["tau"][0]
    assert np.array_equal(ct, g["tau"][1:])
    assert np.array_equal(s.state(), g["u0"])
    assert np.array_equal(s.gradients(), g["grad0"])
    assert np.array_equal(s.edge_fluxes(), g["edgeflux0"])
    assert np.array_equal(s.residual(), g["residual0"])


def test_steps_bitwise(golden_run):
    """States and LIM flags after the reference's run() for each step count."""
    name, case, g = golden_run
    s = oracle.OracleSolver(case)
    steps = sorted(int(k.split("step")[1]) for k in g.files if k.startswith("u_step"))
    done = 0
    for st in steps:
        s.run(st - done)
        done = st
        assert np.array_equal(s.state(), g["u_step%d" % st]), (name, st)
        assert np.array_equal(s.flags(), g["flags_step%d" % st]), (name, st)


def test_bounds_case_exercises_remediation():
    g = gc.load("bounds")
    assert (g["flags_step1"] & 2).sum() > 0
    assert (g["flags_step40"] >> 16).max() >= 0x1e  # counters near the 32-step release


def test_rim_corpus_bitwise():
    g = gc.load("rim")
    out, it = oracle.rim(g["inputs"])
    assert np.array_equal(out, g["outputs"])
    assert it.max() < 50


# ---- textbook exact Riemann solver (Toro, Riemann Solvers, ch. 4) ----------

def _fk(p, rk, pk, ck, g):
    if p > pk:
        A = 2.0 / ((g + 1.0) * rk)
        B = (g - 1.0) / (g + 1.0) * pk
        return (p - pk) * math.sqrt(A / (p + B))
    return 2.0 * ck / (g - 1.0) * ((p / pk) ** ((g - 1.0) / (2.0 * g)) - 1.0)


def exact_star(rl, ul, pl, rr, ur, pr, g=1.4):
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    lo, hi = 1e-14, 1e6
    f = lambda p: _fk(p, rl, pl, cl, g) + _fk(p, rr, pr, cr, g) + ur - ul
    for _ in range(400):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            hi = mid
        else:
            lo = mid
    p = 0.5 * (lo + hi)
    u = 0.5 * (ul + ur) + 0.5 * (_fk(p, rr, pr, cr, g) - _fk(p, rl, pl, cl, g))
    return p, u


def exact_sample0(rl, ul, pl, rr, ur, pr, g=1.4):
    """(rho, u, p) at x/t = 0 (Toro, sec. 4.5)."""
    p, u = exact_star(rl, ul, pl, rr, ur, pr, g)
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    if u >= 0.0:  # left of the contact
        if p > pl:
            sl = ul - cl * math.sqrt((g + 1) / (2 * g) * p / pl + (g - 1) / (2 * g))
            if sl >= 0:
                return rl, ul, pl
            r = rl * ((p / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pl + 1))
            return r, u, p
        shl = ul - cl
        if shl >= 0:
            return rl, ul, pl
        cs = cl * (p / pl) ** ((g - 1) / (2 * g))
        if u - cs < 0:
            return rl * (p / pl) ** (1 / g), u, p
        c = 2 / (g + 1) * (cl + (g - 1) / 2 * ul)
        r = rl * (c / cl) ** (2 / (g - 1))
        return r, c, pl * (c / cl) ** (2 * g / (g - 1))
    if p > pr:
        sr = ur + cr * math.sqrt((g + 1) / (2 * g) * p / pr + (g - 1) / (2 * g))
        if sr <= 0:
            return rr, ur, pr
        r = rr * ((p / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pr + 1))
        return r, u, p
    shr = ur + cr
    if shr <= 0:
        return rr, ur, pr
    cs = cr * (p / pr) ** ((g - 1) / (2 * g))
    if u + cs > 0:
        return rr * (p / pr) ** (1 / g), u, p
    c = 2 / (g + 1) * (cr - (g - 1) / 2 * ur)
    r = rr * (c / cr) ** (2 / (g - 1))
    return r, -c, pr * (c / cr) ** (2 * g / (g - 1))


TORO_STAR = [(0.30313, 0.92745), (0.00189, 0.0), (460.894, 19.5975), (46.0950, -6.19633), (1691.64, 8.68975)]


@pytest.mark.parametrize("k", range(5))
def test_textbook_solver_reproduces_toro_table(k):
    t = gc_toro(k)
    p, u = exact_star(*t)
    ps, us = TORO_STAR[k]
    assert abs(p - ps) <= 5e-4 * abs(ps) + 5e-6  # table printed to 5-6 digits
    assert abs(u - us) <= 5e-4 * max(abs(us), 1.0)


def gc_toro(k):
    from make_golden import TORO
    r = TORO[k]
    return r[0], r[1], r[2], r[3], r[4], r[5]


def test_rim_matches_textbook_sampling():
    """rim_orig's sampled (rho, u, p) at x/t=0 equals the exact solution to the
    accuracy its Newton stop (|residual| <= 1e-5 before the last update) gives."""
    g = gc.load("rim")
    rows, out = g["inputs"], g["outputs"]
    worst = 0.0
    for i in range(0, rows.shape[0], 7):
        rb, pb, ub, vb, re, pe, ue, ve, gam = rows[i]
        r, u, p = exact_sample0(rb, ub, pb, re, ue, pe, gam)
        RI, EI, PI, UI, VI, WI = out[i]
        for a, b in ((RI, r), (UI, u), (PI, p)):
            worst = max(worst, abs(a - b) / max(abs(b), 1e-3))
        # tangential velocity comes from the side of the contact
        assert VI == (vb if u >= 0 else ve) or abs(u) < 1e-6
    assert worst < 1e-6, worst


import sys  # noqa: E402
sys.path.insert(0, gc.GOLDEN)
