"""Host-side logic of the product library, no GPU needed: the C ABI loads and
exports every declared symbol, the host mesh builder and UNV/Triangle
readers reproduce the reference geometry, periodic pairing and batching
match the oracle's, and the task-file errors are reported."""
import hashlib
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import golden_cases as gc
import oracle
from cfd2d import case as cases
from cfd2d import meshgen, native, unv

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "cfd2d.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(cfd2d_[a-z_0-9]+)\s*\(", text)))


def test_header_symbols_exported():
    L = native.lib()
    syms = declared_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(native.SIGNATURES), set(syms) ^ set(native.SIGNATURES)


def test_error_strings():
    L = native.lib()
    assert L.cfd2d_error_string(222) == b"boundary edge not defined in mesh file"
    assert L.cfd2d_error_string(0) == b"ok"


@pytest.mark.parametrize("name", gc.CASES)
def test_mesh_geometry_bitwise(name):
    c = gc.case(name)
    m = native.Mesh.from_trimesh(c.mesh)
    ex = m.export()
    spec = gc.spec(name)
    g = gc.load(name)
    for k in ("cell_geom", "cell_nodes", "cell_edges", "edge_geom", "edge_int"):
        if spec["full"]:
            assert np.array_equal(ex[k], g[k]), k
        else:
            assert hashlib.sha256(np.ascontiguousarray(ex[k]).tobytes()).hexdigest() == spec["geometry_sha256"][k]


@pytest.mark.parametrize("name", ["tiny", "inlet", "bounds"])
def test_unv_reader_matches_arrays(name):
    s = gc.spec(name)
    tm = meshgen.rect_mesh(**s["mesh"])
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "m.unv")
        unv.write_unv(tm, p)
        mr = native.Mesh.read("salome_unv", p)
    ma = native.Mesh.from_trimesh(tm)
    a, b = mr.export(), ma.export()
    for k in ("cell_geom", "cell_nodes", "cell_edges", "edge_geom", "edge_int"):
        assert np.array_equal(a[k], b[k]), k
    # groups come back by name (std::map order); cell groups named as written
    names = mr.group_names(0)
    assert sorted(names) == names
    eg = a["edge_group"]
    for e in np.nonzero(a["edge_int"][:, 3] < 0)[0][:50]:
        gname = names[eg[e]]
        assert gname.startswith("edge_")
    cg = mr.cell_groups()
    cn = mr.group_names(1)
    assert set(cn[i] for i in np.unique(cg)) == {"cells_left", "cells_right"}


def test_unv_reader_errors():
    with tempfile.TemporaryDirectory() as d:
        with pytest.raises(native.Cfd2dError) as e:
            native.Mesh.read("salome_unv", os.path.join(d, "missing.unv"))
        assert e.value.code == 301
        with pytest.raises(native.Cfd2dError) as e:
            native.Mesh.read("gmsh", os.path.join(d, "x"))
        assert e.value.code == 201
        tm = meshgen.rect_mesh(4, 2, seed=1)
        tm.bedges = tm.bedges[:-1]
        tm.bedge_group = tm.bedge_group[:-1]
        p = os.path.join(d, "m.unv")
        unv.write_unv(tm, p)
        with pytest.raises(native.Cfd2dError) as e:
            native.Mesh.read("salome_unv", p)
        assert e.value.code == 222
        text = open(p).read().replace("        41", "        44", 1)
        open(p, "w").write(text)
        with pytest.raises(native.Cfd2dError) as e:
            native.Mesh.read("salome_unv", p)
        assert e.value.code == 221


def test_triangle_reader():
    tm = meshgen.rect_mesh(6, 3, seed=2)
    with tempfile.TemporaryDirectory() as d:
        base = os.path.join(d, "m")
        with open(base + ".node", "w") as f:
            f.write("%d 2 0 0\n" % tm.n_nodes)
            for k, (x, y) in enumerate(tm.xy):
                f.write("%d %.17g %.17g\n" % (k + 1, x, y))
        with open(base + ".ele", "w") as f:
            f.write("%d 3 1\n" % tm.n_cells)
            for k, t in enumerate(tm.tri):
                f.write("%d %d %d %d %d\n" % (k + 1, t[0] + 1, t[1] + 1, t[2] + 1, tm.cell_group[k] + 1))
        with open(base + ".poly", "w") as f:
            f.write("0 2 0 1\n%d 1\n" % tm.n_bedges)
            for k, (a, b) in enumerate(tm.bedges):
                f.write("%d %d %d %d\n" % (k + 1, a + 1, b + 1, tm.bedge_group[k] + 1))
        m = native.Mesh.read("berkeley_triangle", base)
    ref = native.Mesh.from_trimesh(tm).export()
    ex = m.export()
    assert np.array_equal(ex["cell_geom"], ref["cell_geom"])
    assert np.array_equal(ex["edge_geom"], ref["edge_geom"])
    names = m.group_names(0)
    bnd = ex["edge_int"][:, 3] < 0
    got = np.array([int(names[g]) for g in ex["edge_group"][bnd]])
    want = ref["edge_group"][bnd] + 1
    assert np.array_equal(got, want)


def test_periodic_pairing_matches_oracle():
    c = cases.periodic_pulse(12, seed=3)
    m = native.Mesh.from_trimesh(c.mesh, c.periodic_pairs())
    o = oracle.OracleMesh(c.mesh, c.periodic_pairs())
    a, b = m.export(), o.export()
    assert m.ne == o.ne == 3 * 12 * 12
    for k in ("cell_geom", "cell_edges", "edge_geom", "edge_int", "edge_shift"):
        assert np.array_equal(a[k], b[k]), k
    assert (a["edge_int"][:, 3] >= 0).all()


def test_concat_offsets():
    parts = [native.Mesh.from_trimesh(meshgen.bump_channel(ny, seed=k)) for k, ny in enumerate((4, 6, 5))]
    m = native.Mesh.concat(parts)
    co, eo = m.case_offsets()
    assert list(co) == [0, parts[0].nc, parts[0].nc + parts[1].nc, m.nc]
    assert eo[-1] == m.ne == sum(p.ne for p in parts)
    ex = m.export()
    e1 = parts[1].export()
    sl = slice(eo[1], eo[2])
    c2 = e1["edge_int"][:, 3]
    assert np.array_equal(ex["edge_int"][sl, 2], e1["edge_int"][:, 2] + co[1])
    assert np.array_equal(ex["edge_int"][sl, 3], np.where(c2 >= 0, c2 + co[1], -1))
    assert np.array_equal(ex["edge_geom"][sl], e1["edge_geom"])


def test_driver_reports_task_errors():
    exe = os.path.join(REPO, "cfd-2d_amd", "cfd2d", "cfd2d_run")
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "task.xml"), "w").write('<task method="FEM_RKDG"><control/></task>')
        r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and "only FVM_TVD" in r.stdout
        open(os.path.join(d, "task.xml"), "w").write('<task method="FVM_TVD"><control><STEADY value="0"/></control></task>')
        r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and "missing <TAU>" in r.stdout
        assert "missing <TAU>" in open(os.path.join(d, "task.log")).read()
