"""Mass-attenuation tables (NIST XCOM) and their log-log interpolation.

Follows ``deepdrr/material/material.py`` (``Material.get`` 80-95,
``get_coefficients`` 101-110, ``from_string`` 117-155,
``load_material_coeffs_from_lines`` 177-197) and
``deepdrr/material/log_interp.py:6-30``.  The tables under ``data/materials``
are the NIST files deepdrr ships (public physical data, kept byte-identical).
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass
from typing import Dict, List

import numpy as np

_MATERIAL_DIR = os.path.join(os.path.dirname(__file__), "data", "materials")

# Chemical symbols and named materials -> table file (deepdrr/material/mappings.py:1-100).
_ELEMENT_FILES = {}
for _fn in os.listdir(_MATERIAL_DIR):
    _parts = _fn.split("_")
    if len(_parts) == 3 and _parts[0].isdigit():
        _ELEMENT_FILES[_parts[1]] = _fn
NAMED = {
    "soft tissue": "tissue_soft",
    "iron": "26_Fe_Iron",
    "lead": "82_Pb_Lead",
    "copper": "29_Cu_Copper",
    "titanium": "22_Ti_Titanium",
}


@dataclass
class CoefficientEntry:
    energy: float  # MeV
    mu_over_rho: float  # cm^2/g
    mu_en_over_rho: float  # cm^2/g

    def __array__(self, dtype=None, copy=None):
        return np.array([self.energy, self.mu_over_rho, self.mu_en_over_rho], dtype=dtype)


def log_interp(x_interp, x, y):
    """10**interp(log10 x_interp; log10 x, log10 y) in FP64 (np.interp clamps)."""
    xi = np.log10(np.asarray(x_interp, dtype=np.float64))
    return np.power(10.0, np.interp(xi, np.log10(np.asarray(x, np.float64)),
                                    np.log10(np.asarray(y, np.float64))))


class Material:
    _cache: Dict[str, "Material"] = {}
    _custom_map: Dict[str, str] = {}

    def __init__(self, name: str, coefficients: List[CoefficientEntry]):
        self.name = name
        self.coefficients = coefficients

    def __repr__(self):
        return f"Material(name={self.name})"

    def __str__(self):
        return self.name

    def __hash__(self):
        return hash(self.name)

    def __eq__(self, other):
        return isinstance(other, Material) and other.name == self.name

    @property
    def energy(self) -> np.ndarray:
        return np.array([e.energy for e in self.coefficients], dtype=np.float64)

    @property
    def mu_over_rho(self) -> np.ndarray:
        return np.array([e.mu_over_rho for e in self.coefficients], dtype=np.float64)

    @property
    def mu_en_over_rho(self) -> np.ndarray:
        return np.array([e.mu_en_over_rho for e in self.coefficients], dtype=np.float64)

    def _nudged_energy(self) -> np.ndarray:
        # Absorption edges repeat an energy; the first copy is lowered by a
        # relative 1e-9 so np.interp sees a strictly increasing abscissa
        # (material.py:86-92).
        e = self.energy
        for j in range(1, len(e)):
            if e[j] == e[j - 1]:
                e[j - 1] *= 1 - 1e-9
        return e

    def get(self, energy_MeV: float) -> CoefficientEntry:
        e = self._nudged_energy()
        mu = float(log_interp(energy_MeV, e, self.mu_over_rho))
        mu_en = float(log_interp(energy_MeV, e, self.mu_en_over_rho))
        return CoefficientEntry(energy_MeV, mu, mu_en)

    def get_list(self, energies) -> List[CoefficientEntry]:
        return [self.get(e) for e in energies]

    def get_coefficients(self, energy_keV: float) -> CoefficientEntry:
        # keV -> MeV in FP64, as numpy<2 computed it for a float32 input.
        return self.get(float(energy_keV) / 1000.0)

    def mu_over_rho_at_keV(self, energies_keV) -> np.ndarray:
        e = self._nudged_energy()
        return log_interp(np.asarray(energies_keV, np.float64) / 1000.0, e, self.mu_over_rho)

    @classmethod
    def register_map(cls, mapping: Dict[str, str]) -> None:
        cls._custom_map.update(mapping)

    @staticmethod
    def load_material_coeffs_from_lines(lines: List[str]) -> List[CoefficientEntry]:
        out = []
        for line in lines:
            if line.strip() and not line.startswith("_") and not line.startswith("Energy"):
                e, mu, mu_en = map(float, line.strip().split())
                out.append(CoefficientEntry(e, mu, mu_en))
        return out

    @classmethod
    def from_string(cls, name: str, compound_string: bool = False) -> "Material":
        mapped = cls._custom_map.get(name, name)
        mapped = NAMED.get(mapped, mapped)
        mapped = _ELEMENT_FILES.get(mapped, mapped)
        if mapped in cls._cache:
            return cls._cache[mapped]
        if compound_string:
            parts = cls._parse_compound_string(mapped)
            inst = cls(mapped, cls._mix(parts))
            cls._cache[mapped] = inst
            return inst
        path = os.path.join(_MATERIAL_DIR, mapped)
        if not os.path.isfile(path):
            raise AttributeError(f"Material '{mapped}' not found at {path}")
        with open(path) as fh:
            inst = cls(mapped, cls.load_material_coeffs_from_lines(fh.readlines()))
        cls._cache[mapped] = inst
        return inst

    @staticmethod
    def _parse_compound_string(name: str) -> Dict[str, float]:
        matches = re.findall(r"([A-Z][a-z]?)([0-9]+\.[0-9]+)", name)
        if not matches:
            raise ValueError(f"Invalid compound string: {name}")
        parsed = {el: float(fr) for el, fr in matches}
        total = sum(parsed.values())
        if np.isclose(total, 100, atol=1e-2):
            parsed = {el: fr / 100.0 for el, fr in parsed.items()}
            total = sum(parsed.values())
        if not np.isclose(total, 1.0, atol=1e-4):
            raise ValueError(f"Fractions must sum to 1.0, got {total} for {name}")
        return parsed

    @classmethod
    def _mix(cls, parts: Dict[str, float]) -> List[CoefficientEntry]:
        energy = np.array([])
        for el in parts:
            e = cls.from_string(el).energy
            energy = e if energy.size == 0 else np.concatenate([energy, e[~np.isin(e, energy)]])
        energy = np.sort(energy)
        mu = np.zeros_like(energy)
        mu_en = np.zeros_like(energy)
        for el, frac in parts.items():
            ents = cls.from_string(el).get_list(energy)
            mu = mu + np.array([c.mu_over_rho for c in ents]) * frac
            mu_en = mu_en + np.array([c.mu_en_over_rho for c in ents]) * frac
        return [CoefficientEntry(e, mu[i], mu_en[i]) for i, e in enumerate(energy)]


def attenuation_table(material_names: List[str], energies_keV_f32: np.ndarray) -> np.ndarray:
    """Flat ``[bin * M + m]`` float32 mu/rho table (projector.py:1676-1689)."""
    e = np.asarray(energies_keV_f32, dtype=np.float32)
    out = np.zeros(len(e) * len(material_names), dtype=np.float32)
    for m, name in enumerate(material_names):
        mat = Material.from_string(name)
        for b in range(len(e)):
            out[b * len(material_names) + m] = np.float32(mat.get_coefficients(e[b]).mu_over_rho)
    return out
