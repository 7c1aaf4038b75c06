"""Material / Materials: the reference's parameter loading.

Mirrors ``uepm::pseudopotential::Material`` and ``Materials``
(src/EPP/Material.h:14-187, src/EPP/Material.cpp:33-71) including their unit
conversions: form factors and non-local alpha / A2 from Ry to eV, radii from
Angstrom to metres, spin-orbit parameters as read.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import yaml

from ._abi import uepm_material

# src/Common/physical_constants.hpp:18-38 (same operation order)
_PI = 3.14159265358979323846
_H = 6.62607015e-34
_QE = 1.602176634e-19
_HARTREE_J = 4.3597447222071e-18
RYD_TO_EV = (_HARTREE_J / _QE) / 2.0
ANGSTROM_TO_M = 1e-10

PARAMS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "params")
DEFAULT_MATERIALS_FILE = os.path.join(PARAMS_DIR, "materials.yaml")


@dataclass
class NonLocalParameters:
    """src/EPP/NonLocalParameters.hpp:32-131 (after populate: eV, 1, m)."""

    alpha_0_cation: float = 0.0
    beta_0_cation: float = 0.0
    A2_cation: float = 0.0
    alpha_0_anion: float = 0.0
    beta_0_anion: float = 0.0
    A2_anion: float = 0.0
    R0_cation: float = 0.0
    R2_cation: float = 0.0
    R0_anion: float = 0.0
    R2_anion: float = 0.0
    well_type: str = "unknown"


@dataclass
class SpinOrbitParameters:
    """src/EPP/SpinOrbitParameters.hpp:27-101 (raw YAML values)."""

    period_anion: int = 0
    period_cation: int = 0
    radial_extent_anion: float = 0.0
    radial_extent_cation: float = 0.0
    alpha: float = 0.0
    mu: float = 0.0


@dataclass
class Material:
    name: str
    lattice_constant: float  # Angstrom (Material::m_lattice_constant)
    V_S: tuple  # eV, |G|^2 = 3, 4, 8, 11
    V_A: tuple
    non_local: NonLocalParameters = field(default_factory=NonLocalParameters)
    spin_orbit: SpinOrbitParameters = field(default_factory=SpinOrbitParameters)
    is_non_local_parameters_populated: bool = False
    is_spin_orbit_parameters_populated: bool = False

    def get_name(self) -> str:
        return self.name

    def get_lattice_constant_meter(self) -> float:
        return ANGSTROM_TO_M * self.lattice_constant

    def to_c(self) -> uepm_material:
        m = uepm_material()
        m.lattice_constant_angstrom = self.lattice_constant
        for i in range(4):
            m.V_S[i] = self.V_S[i]
            m.V_A[i] = self.V_A[i]
        nl = self.non_local
        m.nl_populated = int(self.is_non_local_parameters_populated)
        m.nl_well_type = {"square": 0, "gaussian": 1}.get(nl.well_type, 2)
        for f in ("alpha_0_cation", "beta_0_cation", "A2_cation", "alpha_0_anion", "beta_0_anion", "A2_anion",
                  "R0_cation", "R2_cation", "R0_anion", "R2_anion"):
            setattr(m, f, getattr(nl, f))
        so = self.spin_orbit
        m.soc_populated = int(self.is_spin_orbit_parameters_populated)
        m.period_anion = so.period_anion
        m.period_cation = so.period_cation
        m.radial_extent_anion = so.radial_extent_anion
        m.radial_extent_cation = so.radial_extent_cation
        m.alpha_soc = so.alpha
        m.mu_soc = so.mu
        return m


class Materials:
    """Materials::load_material_parameters (src/EPP/Material.cpp:33-71)."""

    def __init__(self) -> None:
        self.materials: dict[str, Material] = {}

    def load_material_parameters(self, filename: str = DEFAULT_MATERIALS_FILE) -> None:
        with open(filename) as fh:
            doc = yaml.safe_load(fh)
        for node in doc["materials"]:
            symbol = str(node["symbol"])
            a = float(node["lattice-constant"])
            pp = node["pseudo-potential-parameters"]
            r = RYD_TO_EV
            V3S, V8S, V11S = r * pp["V3S"], r * pp["V8S"], r * pp["V11S"]
            V3A, V4A, V11A = r * pp["V3A"], r * pp["V4A"], r * pp["V11A"]
            V4S = r * pp["V4S"] if "V4S" in pp else 0.0
            V8A = r * pp["V8A"] if "V8A" in pp else 0.0
            mat = Material(symbol, a, (V3S, V4S, V8S, V11S), (V3A, V4A, V8A, V11A))
            nl = node.get("non-local-parameters")
            if nl:
                # NonLocalParameters.hpp:91-116
                wt = str(nl["well_type"])
                if wt not in ("square", "gaussian"):
                    raise RuntimeError("Unknown non-local well type")
                mat.non_local = NonLocalParameters(
                    alpha_0_cation=r * nl["alpha_0_cation"], beta_0_cation=float(nl["beta_0_cation"]),
                    A2_cation=r * nl["A2_cation"], alpha_0_anion=r * nl["alpha_0_anion"],
                    beta_0_anion=float(nl["beta_0_anion"]), A2_anion=r * nl["A2_anion"],
                    R0_cation=float(nl["R0_cation"]) * ANGSTROM_TO_M, R2_cation=float(nl["R2_cation"]) * ANGSTROM_TO_M,
                    R0_anion=float(nl["R0_anion"]) * ANGSTROM_TO_M, R2_anion=float(nl["R2_anion"]) * ANGSTROM_TO_M,
                    well_type=wt)
                mat.is_non_local_parameters_populated = True
            so = node.get("spin-orbit-parameters")
            if so:
                # SpinOrbitParameters.hpp:80-87
                mat.spin_orbit = SpinOrbitParameters(
                    period_anion=int(so["period_anion"]), period_cation=int(so["period_cation"]),
                    radial_extent_anion=float(so["radial_extent_anion"]),
                    radial_extent_cation=float(so["radial_extent_cation"]),
                    alpha=float(so["alpha_soc"]), mu=float(so["mu_soc"]))
                mat.is_spin_orbit_parameters_populated = True
            self.materials[symbol] = mat


def load(symbol: str, filename: str = DEFAULT_MATERIALS_FILE) -> Material:
    ms = Materials()
    ms.load_material_parameters(filename)
    return ms.materials[symbol]
