"""Solver configuration shared by the Python host layer.

`SolverConfig` carries what the reference reads from <control>, <limits>,
<materials> and <boundaries> of a FVM_TVD task file
(src/methods/fvm_tvd.cpp:27-121, src/bnd_cond.cpp:10-72) plus the optional
extension switches.  Defaults are the reference behaviour: exact Godunov
flux with GAM = 1.4 (fvm_tvd.cpp:345), Green-Gauss gradients, no limiter,
Heun RK2, no viscous terms, area-based time step.
"""
from __future__ import annotations

from dataclasses import dataclass, field
import math

import numpy as np

FLUX = {"GODUNOV": 0, "LAX": 1, "HLLC": 2, "ROE": 3}
GRADIENT = {"GREEN_GAUSS": 0, "GG": 0, "LSQ": 1, "LEAST_SQUARES": 1}
LIMITER = {"NONE": 0, "MINMOD": 1}
CFL_MODE = {"REFERENCE": 0, "SPECTRAL": 1}
BND = {"BOUND_INLET": 1, "BOUND_OUTLET": 2, "BOUND_WALL_SLIP": 3, "BOUND_WALL_NO_SLIP": 3,
       "BOUND_PRESSURE_OUTLET": 4}

GAS_R = 8.314472  # src/global.cpp:6

MAXMAT = 8
MAXBC = 64


@dataclass
class Material:
    name: str = "air"
    M: float = GAS_R      # molar mass: M = gR gives a gas constant of 1
    Cp: float = 3.5       # Cp = 3.5 with R = 1 gives gamma = 1.4
    K: float = 0.0        # conductivity (extension; 0 -> mu*Cp/Pr)
    ML: float = 0.0       # dynamic viscosity (extension)

    @property
    def Cv(self):
        return self.Cp - GAS_R / self.M

    @property
    def gamma(self):
        return self.Cp / self.Cv

    @property
    def R(self):
        return GAS_R / self.M


@dataclass
class Boundary:
    name: str
    type: str
    Vx: float = 0.0
    Vy: float = 0.0
    T: float = 0.0
    P: float = 0.0
    pair: str = ""
    shift: tuple = (0.0, 0.0)

    @property
    def code(self):
        return BND[self.type]

    def params(self):
        if self.type == "BOUND_PRESSURE_OUTLET":
            return (self.P, 0.0, 0.0, 0.0)
        return (self.Vx, self.Vy, self.T, self.P)


@dataclass
class Region:
    name: str
    Vx: float
    Vy: float
    T: float
    P: float
    material: int = 0
    cell_type: int = 0


@dataclass
class SolverConfig:
    flux: str = "GODUNOV"
    gradient: str = "GREEN_GAUSS"
    limiter: str = "NONE"
    rk_stages: int = 2
    viscous: bool = False
    steady: bool = False
    cfl_mode: str = "REFERENCE"
    max_newton: int = 200
    cfl: float = 0.5
    tau: float = 1.0
    tmax: float = 1.0e300
    step_max: int = 1000
    gam_flux: float = 1.4
    ro_min: float = -math.inf
    ro_max: float = math.inf
    p_min: float = -math.inf
    p_max: float = math.inf
    u_max: float = math.inf
    prandtl: float = 0.72
    materials: list = field(default_factory=lambda: [Material()])
    boundaries: list = field(default_factory=list)
    regions: list = field(default_factory=list)

    def fill(self, st):
        """Copy into a ctypes structure with the cfd2d_params / or_cfg layout."""
        st.flux = FLUX[self.flux]
        st.gradient = GRADIENT[self.gradient]
        st.limiter = LIMITER[self.limiter]
        st.rk_stages = int(self.rk_stages)
        st.viscous = int(bool(self.viscous))
        st.steady = int(bool(self.steady))
        st.cfl_mode = CFL_MODE[self.cfl_mode]
        st.max_newton = int(self.max_newton)
        st.cfl = self.cfl
        st.tau = self.tau
        st.gam_flux = self.gam_flux
        st.ro_min = self.ro_min
        st.ro_max = self.ro_max
        st.p_min = self.p_min
        st.p_max = self.p_max
        st.u_max = self.u_max
        st.prandtl = self.prandtl
        if len(self.materials) > MAXMAT:
            raise ValueError("at most %d materials" % MAXMAT)
        st.nmat = len(self.materials)
        for i, m in enumerate(self.materials):
            st.M[i] = m.M
            st.Cp[i] = m.Cp
            st.K[i] = m.K
            st.ML[i] = m.ML
        if not hasattr(st, "bc_type"):
            return st  # the library takes boundary objects as separate arrays
        bcs = [b for b in self.boundaries if b.type != "BOUND_PERIODIC"]
        if len(bcs) > MAXBC:
            raise ValueError("at most %d boundary conditions" % MAXBC)
        st.nbc = len(bcs)
        for i, b in enumerate(bcs):
            st.bc_type[i] = b.code
            for k, v in enumerate(b.params()):
                st.bc_par[i][k] = v
        return st

    def bc_index(self, name):
        k = 0
        for b in self.boundaries:
            if b.type == "BOUND_PERIODIC":
                continue
            if b.name == name:
                return k
            k += 1
        return -1


def region_state(region: Region, mat: Material):
    """Conserved state of a region: URS(2) then URS(1) (fvm_tvd.cpp:91-92), then
    convertParToCons (fvm_tvd.cpp:795-801)."""
    p = region.P
    T = region.T
    r = p * mat.M / (T * GAS_R)
    gam = mat.Cp / (mat.Cp - GAS_R / mat.M)
    e = p / (r * (gam - 1))
    u = region.Vx
    v = region.Vy
    return np.array([r, r * u, r * v, r * (e + 0.5 * (u * u + v * v))])


def prim_to_cons(rho, u, v, p, gamma):
    """Conserved (rho, rho u, rho v, rho E) from primitives with e = p/(rho (gamma-1))."""
    rho = np.asarray(rho, dtype=np.float64)
    e = p / (rho * (gamma - 1.0))
    return np.stack([rho, rho * u, rho * v, rho * (e + 0.5 * (u * u + v * v))])
