"""Writes the golden fixtures of tests/golden/ from the compiled reference.

Run in a tree that has the reference sources (the harness is built by
``make -C oracle ref``).  Every case regenerates its mesh from the project's
seeded generator, writes the UNV file and task.xml the reference reads, runs
oracle/_ref/ref_harness under a time limit and stores the arrays it dumps in
one compressed .npz per case plus manifest.json (sizes, seeds, sha256 of the
UNV text, so the tests can show that the regenerated mesh is the same file).

Cases (all reference-exact: Green-Gauss, exact Godunov, Heun RK2):
  tiny     10x2  rectangle, mild shock tube, slip walls + outlets
  inlet    50x5  rectangle, inlet/outlet/slip walls, non-zero velocity
  steady   30x4  rectangle, STEADY=1 (local time steps), inlet
  bounds   50x5  rectangle, tight |u| limit -> LIM flags + remediation
  rect8k   200x20 rectangle (8000 cells), mild shock tube, 1000 steps
  twomat   50x5  rectangle, inlet/outlet/slip walls, two materials (the
           right half a gas with R = 0.5 and gamma = 4/3): the per-cell
           material of the EOS, the ghosts and the time step
  rim      rim_orig on Toro's five tests and 4000 seeded random pairs

`python3 make_golden.py twomat` regenerates only the named cases and keeps
the others' manifest entries.

The full Sod state (rho,p = 1,1 | 0.125,0.1) is not a case: with the
reference's unlimited reconstruction it produces negative reconstructed
pressures in the first step, and rim_orig's uncapped Newton iteration
(src/global.cpp:278-307) then never terminates.
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "cfd-2d_amd"))

from cfd2d import meshgen, unv  # noqa: E402

HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_harness")

MILD = dict(left=dict(Vx=0.0, Vy=0.0, T=1.0, P=1.0), right=dict(Vx=0.0, Vy=0.0, T=1.0, P=0.5))


def case_specs():
    walls_outlets = [dict(name="edge_bottom", type="BOUND_WALL_SLIP"),
                     dict(name="edge_top", type="BOUND_WALL_NO_SLIP"),
                     dict(name="edge_left", type="BOUND_OUTLET"),
                     dict(name="edge_right", type="BOUND_OUTLET")]
    inlet = [dict(name="edge_bottom", type="BOUND_WALL_SLIP"),
             dict(name="edge_top", type="BOUND_WALL_SLIP"),
             dict(name="edge_left", type="BOUND_INLET", Vx=0.3, Vy=0.05, T=1.0, P=1.0),
             dict(name="edge_right", type="BOUND_OUTLET")]
    regions_mild = [dict(name="cells_left", **MILD["left"]), dict(name="cells_right", **MILD["right"])]
    regions_flow = [dict(name="cells_left", Vx=0.3, Vy=0.0, T=1.0, P=1.0),
                    dict(name="cells_right", Vx=0.2, Vy=0.0, T=1.1, P=0.8)]
    return [
        dict(name="tiny", mesh=dict(nx=10, ny=2, lx=1.0, ly=0.2, seed=5), regions=regions_mild,
             boundaries=walls_outlets, steady=0, cfl=0.5, steps=[1, 10, 100], full=True),
        dict(name="inlet", mesh=dict(nx=50, ny=5, lx=1.0, ly=0.2, seed=6), regions=regions_flow,
             boundaries=inlet, steady=0, cfl=0.5, steps=[1, 10, 100], full=True),
        dict(name="steady", mesh=dict(nx=30, ny=4, lx=1.0, ly=0.2, seed=8), regions=regions_flow,
             boundaries=inlet, steady=1, cfl=0.4, steps=[1, 10, 50], full=True),
        dict(name="bounds", mesh=dict(nx=50, ny=5, lx=1.0, ly=0.2, seed=9), regions=regions_mild,
             boundaries=walls_outlets, steady=0, cfl=0.5, steps=[1, 5, 20, 40, 60],
             limits=dict(ro_min=1.0e-6, ro_max=10.0, p_min=1.0e-6, p_max=10.0, u_max=0.004),
             full=True),
        dict(name="rect8k", mesh=dict(nx=200, ny=20, lx=1.0, ly=0.2, seed=1), regions=regions_mild,
             boundaries=walls_outlets, steady=0, cfl=0.5, steps=[1, 10, 100, 1000], full=False),
        dict(name="twomat", mesh=dict(nx=50, ny=5, lx=1.0, ly=0.2, seed=12),
             materials=[dict(name="air", M=8.314472, Cp=3.5, K=0.0, ML=0.0),
                        dict(name="heavy", M=16.628944, Cp=2.0, K=0.0, ML=0.0)],
             regions=[dict(name="cells_left", material=0, Vx=0.3, Vy=0.0, T=1.0, P=1.0),
                      dict(name="cells_right", material=1, Vx=0.2, Vy=0.0, T=1.4, P=0.8)],
             boundaries=inlet, steady=0, cfl=0.5, steps=[1, 10, 100], full=True),
    ]


def build_case_inputs(spec, workdir):
    m = meshgen.rect_mesh(**spec["mesh"])
    mesh_path = os.path.join(workdir, "mesh.unv")
    unv.write_unv(m, mesh_path)
    text = unv.task_xml("mesh.unv", regions=spec["regions"], boundaries=spec["boundaries"],
                        materials=spec.get("materials"), steady=spec["steady"], cfl=spec["cfl"],
                        step_max=max(spec["steps"]), limits=spec.get("limits"))
    with open(os.path.join(workdir, "task.xml"), "w") as f:
        f.write(text)
    with open(mesh_path, "rb") as f:
        sha = hashlib.sha256(f.read()).hexdigest()
    return m, text, sha


def _load(d, name, dtype):
    return np.fromfile(os.path.join(d, name), dtype=dtype)


def run_case(spec, outdir, timeout=600):
    with tempfile.TemporaryDirectory() as wd:
        m, text, sha = build_case_inputs(spec, wd)
        cmd = [HARNESS, "golden", "task.xml", "."] + [str(s) for s in spec["steps"]]
        subprocess.run(cmd, cwd=wd, check=True, timeout=timeout, stdout=subprocess.DEVNULL)
        nn, nc, ne = (int(v) for v in _load(wd, "counts.i32", np.int32))
        arrs = dict(
            counts=np.array([nn, nc, ne], np.int32),
            tau=_load(wd, "tau.f64", np.float64),
            u0=_load(wd, "u0.f64", np.float64).reshape(4, nc),
            grad0=_load(wd, "grad0.f64", np.float64).reshape(8, nc),
            edgeflux0=_load(wd, "edgeflux0.f64", np.float64).reshape(ne, 4),
            residual0=_load(wd, "residual0.f64", np.float64).reshape(4, nc),
            residual0_scale=_load(wd, "residual0_scale.f64", np.float64).reshape(4, nc),
        )
        geo = dict(
            cell_geom=_load(wd, "cell_geom.f64", np.float64).reshape(nc, 4),
            cell_nodes=_load(wd, "cell_nodes.i32", np.int32).reshape(nc, 3),
            cell_edges=_load(wd, "cell_edges.i32", np.int32).reshape(nc, 3),
            edge_geom=_load(wd, "edge_geom.f64", np.float64).reshape(ne, 9),
            edge_int=_load(wd, "edge_int.i32", np.int32).reshape(ne, 5),
        )
        if spec["full"]:
            arrs.update(geo)
            arrs["recon0"] = _load(wd, "recon0.f64", np.float64).reshape(ne, 2, 8)
            arrs["gpflux0"] = _load(wd, "gpflux0.f64", np.float64).reshape(ne, 2, 4)
        geo_sha = {k: hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest() for k, v in geo.items()}
        for s in spec["steps"]:
            arrs["u_step%d" % s] = _load(wd, "u_step%d.f64" % s, np.float64).reshape(4, nc)
            arrs["flags_step%d" % s] = _load(wd, "flags_step%d.u32" % s, np.uint32)
        np.savez_compressed(os.path.join(outdir, spec["name"] + ".npz"), **arrs)
        out = dict(name=spec["name"], mesh=spec["mesh"], regions=spec["regions"],
                    boundaries=spec["boundaries"], steady=spec["steady"], cfl=spec["cfl"],
                    steps=spec["steps"], limits=spec.get("limits"), unv_sha256=sha,
                    geometry_sha256=geo_sha, counts=[nn, nc, ne], full=spec["full"],
                    task_sha256=hashlib.sha256(text.encode()).hexdigest())
        if spec.get("materials"):
            out["materials"] = spec["materials"]
        return out


TORO = [  # (rhoL, uL, pL, rhoR, uR, pR) -- Toro, Riemann Solvers, Table 4.1
    (1.0, 0.0, 1.0, 0.125, 0.0, 0.1),
    (1.0, -2.0, 0.4, 1.0, 2.0, 0.4),
    (1.0, 0.0, 1000.0, 1.0, 0.0, 0.01),
    (1.0, 0.0, 0.01, 1.0, 0.0, 100.0),
    (5.99924, 19.5975, 460.894, 5.99242, -6.19633, 46.0950),
]


def rim_corpus(n=4000, seed=11):
    idx = np.arange(n)
    u = [0.5 * (meshgen.uniform(seed, 20 + k, idx) + 1.0) for k in range(8)]
    rb = 0.05 + 2.0 * u[0]
    pb = 0.05 + 2.0 * u[1]
    ub = 2.0 * (u[2] - 0.5)
    vb = 2.0 * (u[3] - 0.5)
    re = 0.05 + 2.0 * u[4]
    pe = 0.05 + 2.0 * u[5]
    ue = 2.0 * (u[6] - 0.5)
    ve = 2.0 * (u[7] - 0.5)
    rows = [np.stack([rb, pb, ub, vb, re, pe, ue, ve, np.full(n, 1.4)], axis=1)]
    toro = np.array([[t[0], t[2], t[1], 0.0, t[3], t[5], t[4], 0.0, 1.4] for t in TORO])
    return np.concatenate([toro] + rows)


def run_rim(outdir):
    rows = rim_corpus()
    with tempfile.TemporaryDirectory() as wd:
        inp = os.path.join(wd, "in.bin")
        out = os.path.join(wd, "out.bin")
        rows.astype("<f8").tofile(inp)
        subprocess.run([HARNESS, "rim", inp, out], check=True, timeout=120)
        res = np.fromfile(out, dtype="<f8").reshape(-1, 6)
    np.savez_compressed(os.path.join(outdir, "rim.npz"), inputs=rows, outputs=res)
    return dict(name="rim", rows=int(rows.shape[0]), toro_rows=len(TORO))


def main():
    if not os.path.exists(HARNESS):
        sys.exit("reference harness missing: run `make -C oracle ref` where the reference tree exists")
    only = set(sys.argv[1:])
    manifest = dict(cases=[], generator="tests/golden/make_golden.py",
                    harness="oracle/_ref/ref_harness (reference sources, g++ -O2 -ffp-contract=off)")
    old = {}
    if only:
        with open(os.path.join(HERE, "manifest.json")) as f:
            prev = json.load(f)
        old = {c["name"]: c for c in prev["cases"]}
        manifest["rim"] = prev["rim"]
    for spec in case_specs():
        if only and spec["name"] not in only:
            manifest["cases"].append(old[spec["name"]])
            continue
        print("case", spec["name"], flush=True)
        manifest["cases"].append(run_case(spec, HERE))
    if not only or "rim" in only:
        manifest["rim"] = run_rim(HERE)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)
    print("done")


if __name__ == "__main__":
    main()
