"""Rebuilds the golden cases of tests/golden/manifest.json as cfd2d Cases."""
import json
import math
import os

import numpy as np

from cfd2d import meshgen
from cfd2d.case import Case
from cfd2d.config import Boundary, Material, Region, SolverConfig

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


def spec(name):
    for c in manifest()["cases"]:
        if c["name"] == name:
            return c
    raise KeyError(name)


def load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def case(name, **over):
    s = spec(name)
    m = meshgen.rect_mesh(**s["mesh"])
    regions = [Region(r["name"], r["Vx"], r["Vy"], r["T"], r["P"], material=r.get("material", 0))
               for r in s["regions"]]
    bnds = [Boundary(b["name"], b["type"], Vx=b.get("Vx", 0.0), Vy=b.get("Vy", 0.0), T=b.get("T", 0.0),
                     P=b.get("P", 0.0)) for b in s["boundaries"]]
    lim = s.get("limits") or dict(ro_min=1.0e-10, ro_max=1.0e10, p_min=1.0e-10, p_max=1.0e10, u_max=1.0e10)
    kw = dict(regions=regions, boundaries=bnds, steady=bool(s["steady"]), cfl=s["cfl"], tau=1.0,
              ro_min=lim["ro_min"], ro_max=lim["ro_max"], p_min=lim["p_min"], p_max=lim["p_max"],
              u_max=lim["u_max"], max_newton=500)
    if s.get("materials"):
        kw["materials"] = [Material(name=x["name"], M=x["M"], Cp=x["Cp"], K=x["K"], ML=x["ML"])
                           for x in s["materials"]]
    kw.update(over)
    return Case(m, SolverConfig(**kw), name=name)


CASES = ["tiny", "inlet", "steady", "bounds", "rect8k", "twomat"]
FULL = ["tiny", "inlet", "steady", "bounds", "twomat"]
