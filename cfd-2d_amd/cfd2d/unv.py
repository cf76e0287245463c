"""Salome UNV mesh writer and FVM_TVD task-file writer.

The UNV writer emits exactly what the reference's UNV reader accepts
(src/mesh/MeshReaderSalomeUnv.cpp:267-286 block delimiting, 313-334 nodes,
337-390 elements, 393-425 groups): node labels 1..V in order, one type-11
element per boundary edge, one type-41 element per triangle, and one 2467
group per named edge or cell group.  Element labels are compact (edges first,
then triangles), so the reader's dense label table stays small.
"""
from __future__ import annotations

import io

import numpy as np

from .meshgen import TriMesh


def _fmt_ints(vals):
    return "".join("%10d" % v for v in vals) + "\n"


def write_unv(mesh: TriMesh, path: str) -> None:
    out = io.StringIO()
    w = out.write
    # nodes (dataset 2411): record 1 = label + 3 ints, record 2 = coordinates
    w("    -1\n  2411\n")
    xy = mesh.xy
    lines = []
    for k in range(mesh.n_nodes):
        lines.append("%10d%10d%10d%10d\n%25.16E%25.16E%25.16E\n" % (k + 1, 1, 1, 11, xy[k, 0], xy[k, 1], 0.0))
    w("".join(lines))
    w("    -1\n")
    # elements (dataset 2412)
    w("    -1\n  2412\n")
    lines = []
    nb = mesh.n_bedges
    for k in range(nb):
        a, b = mesh.bedges[k]
        lines.append("%10d%10d%10d%10d%10d%10d\n%10d%10d%10d\n%10d%10d\n"
                     % (k + 1, 11, 2, 1, 7, 2, 0, 0, 0, a + 1, b + 1))
    for c in range(mesh.n_cells):
        a, b, d = mesh.tri[c]
        lines.append("%10d%10d%10d%10d%10d%10d\n%10d%10d%10d\n"
                     % (nb + c + 1, 41, 2, 1, 7, 3, a + 1, b + 1, d + 1))
    w("".join(lines))
    w("    -1\n")
    # groups (dataset 2467): entity type 8 = element, tag = element label
    w("    -1\n  2467\n")
    gid = 1
    groups = []
    for g, name in enumerate(mesh.edge_group_names):
        groups.append((name, np.nonzero(mesh.bedge_group == g)[0] + 1))
    for g, name in enumerate(mesh.cell_group_names):
        groups.append((name, np.nonzero(mesh.cell_group == g)[0] + nb + 1))
    for name, labels in groups:
        n = len(labels)
        if n == 0:
            continue
        w(_fmt_ints([gid, 0, 0, 0, 0, 0, 0, n]))
        w(name + "\n")
        lines = []
        for i in range(0, n - 1, 2):
            lines.append(_fmt_ints([8, labels[i], 0, 0, 8, labels[i + 1], 0, 0]))
        if n % 2 == 1:
            lines.append(_fmt_ints([8, labels[n - 1], 0, 0]))
        w("".join(lines))
        gid += 1
    w("    -1\n")
    with open(path, "w") as f:
        f.write(out.getvalue())


def _xml_escape(s):
    return str(s).replace("&", "&amp;").replace("<", "&lt;").replace(">", "&gt;")


def task_xml(mesh_path, *, regions, boundaries, materials=None, steady=0, tau=1.0, tmax=1.0e9,
             cfl=0.5, step_max=1000, file_output_step=1000000, log_output_step=1000000,
             limits=None, mesh_type="salome_unv", extensions=None, method="FVM_TVD", initial_file=None):
    """Text of a task.xml in the reference's FVM_TVD schema.

    regions: list of dict(name, material=0, Vx, Vy, T, P, cell_type=0)
    boundaries: list of dict(name, type, edge_type=0, Vx/Vy/T/P for inlets,
                pair/shift for the periodic extension)
    extensions: optional dict of extension control nodes written as
                <NAME value="..."/> inside <control> (FLUX, GRADIENT, LIMITER, ...)
    initial_file: optional <initial file="..."/> (extension): raw little-endian
                float64 conserved state, double[4][cells] in the mesh file's
                cell order, replacing the regions' uniform states
    """
    if materials is None:
        materials = [dict(name="air", M=8.314472, Cp=3.5, K=0.0, ML=0.0)]
    if limits is None:
        limits = dict(ro_min=1.0e-10, ro_max=1.0e10, p_min=1.0e-10, p_max=1.0e10, u_max=1.0e10)
    s = io.StringIO()
    w = s.write
    w('<?xml version="1.0" encoding="utf-8"?>\n')
    w('<task method="%s">\n' % method)
    w("  <control>\n")
    w('    <STEADY value="%d"/>\n' % steady)
    w('    <TAU value="%.17g"/>\n' % tau)
    w('    <TMAX value="%.17g"/>\n' % tmax)
    w('    <CFL value="%.17g"/>\n' % cfl)
    w('    <STEP_MAX value="%d"/>\n' % step_max)
    w('    <FILE_OUTPUT_STEP value="%d"/>\n' % file_output_step)
    w('    <LOG_OUTPUT_STEP value="%d"/>\n' % log_output_step)
    for k, v in (extensions or {}).items():
        w('    <%s value="%s"/>\n' % (k, _xml_escape(v)))
    w("  </control>\n")
    w("  <limits>\n")
    w('    <ro min="%.17g" max="%.17g"/>\n' % (limits["ro_min"], limits["ro_max"]))
    w('    <p min="%.17g" max="%.17g"/>\n' % (limits["p_min"], limits["p_max"]))
    w('    <u max="%.17g"/>\n' % limits["u_max"])
    w("  </limits>\n")
    w('  <materials count="%d">\n' % len(materials))
    for i, m in enumerate(materials):
        w('    <material id="%d">\n      <name>%s</name>\n      <parameters>\n' % (i, _xml_escape(m["name"])))
        for key in ("M", "Cp", "K", "ML"):
            w('        <%s value="%.17g"/>\n' % (key, m[key]))
        w("      </parameters>\n    </material>\n")
    w("  </materials>\n")
    w('  <regions count="%d">\n' % len(regions))
    for i, r in enumerate(regions):
        w('    <region id="%d">\n' % i)
        w('      <material id="%d"/>\n' % r.get("material", 0))
        w('      <cell type="%d"/>\n' % r.get("cell_type", 0))
        w("      <name>%s</name>\n      <parameters>\n" % _xml_escape(r["name"]))
        for key in ("Vx", "Vy", "T", "P"):
            w('        <%s value="%.17g"/>\n' % (key, r[key]))
        w("      </parameters>\n    </region>\n")
    w("  </regions>\n")
    w("  <boundaries>\n")
    for b in boundaries:
        w('    <boundCond edgeType="%d">\n' % b.get("edge_type", 0))
        w("      <type>%s</type>\n" % b["type"])
        w("      <name>%s</name>\n" % _xml_escape(b["name"]))
        if b["type"] == "BOUND_INLET":
            w("      <parameters>\n")
            for key in ("Vx", "Vy", "T", "P"):
                w('        <%s value="%.17g"/>\n' % (key, b[key]))
            w("      </parameters>\n")
        if b["type"] == "BOUND_PRESSURE_OUTLET":
            w('      <parameters>\n        <P value="%.17g"/>\n      </parameters>\n' % b["P"])
        if b["type"] == "BOUND_PERIODIC":
            w("      <pair>%s</pair>\n" % _xml_escape(b["pair"]))
            w('      <shift x="%.17g" y="%.17g"/>\n' % tuple(b["shift"]))
        w("    </boundCond>\n")
    w("  </boundaries>\n")
    w("  <mesh>\n")
    w('    <name value="%s"/>\n' % _xml_escape(mesh_path))
    w('    <filesType value="%s"/>\n' % mesh_type)
    w("  </mesh>\n")
    if initial_file is not None:
        w('  <initial file="%s"/>\n' % _xml_escape(initial_file))
    w("</task>\n")
    return s.getvalue()


def sod_task(mesh_path, *, step_max=1000, cfl=0.5, limits=None, extensions=None, tau=1.0):
    """Task for the Sod-type rectangle: (rho,p) = (1,1) left and (0.125,0.1) right.

    The material has M = gR and Cp = 3.5, so the gas constant is 1, gamma is
    1.4 and T = p / rho.
    """
    regions = [dict(name="cells_left", Vx=0.0, Vy=0.0, T=1.0, P=1.0),
               dict(name="cells_right", Vx=0.0, Vy=0.0, T=0.8, P=0.1)]
    boundaries = [dict(name="edge_bottom", type="BOUND_WALL_SLIP"),
                  dict(name="edge_top", type="BOUND_WALL_SLIP"),
                  dict(name="edge_left", type="BOUND_OUTLET"),
                  dict(name="edge_right", type="BOUND_OUTLET")]
    return task_xml(mesh_path, regions=regions, boundaries=boundaries, cfl=cfl,
                    step_max=step_max, limits=limits, extensions=extensions, tau=tau)
