"""ASE calculator over the am1g GPU engine: the molecular path of ``am1_rs.ase.AM1``
(python/am1_rs/ase.py:354-379). ASE units throughout: energies eV, forces eV/A, positions A,
dipole e*A; ``field`` in V/A (eV per e*A), converted at this boundary to the atomic units of
``am1g.gradient`` / ``am1g.single_point``.

Charge, multiplicity, reference, method and field come from the constructor or, per structure,
from ``atoms.info`` (the same override rule as the reference). Periodic cells are not part of
the GPU engine and raise ``NotImplementedError``.
"""
from __future__ import annotations

import numpy as np

try:
    from ase.calculators.calculator import Calculator, all_changes
    from ase.units import Debye
except ImportError as exc:  # pragma: no cover
    raise ImportError("The am1g ASE calculator requires ASE (pip install ase).") from exc

import am1g

_BOHR = 0.529167  # Angstrom (constants.rs:10)
_HARTREE = 27.21  # eV (constants.rs:13)


class AM1(Calculator):
    """AM1/RM1 calculator on the GPU (eV, eV/A, A). Populates ``energy``, ``free_energy``,
    ``forces`` (from the same SCF as the energy when forces are asked for), ``charges``
    (Mulliken, e) and ``dipole`` (e*A)."""

    implemented_properties = ["energy", "free_energy", "forces", "charges", "dipole"]
    default_parameters = {"charge": 0.0, "multiplicity": 1, "reference": "auto", "method": "am1", "field": None}

    def __init__(self, **kwargs):
        super().__init__(**kwargs)
        self._last_state = None

    def _field_au(self, atoms):
        field = self.parameters.get("field")
        info = getattr(atoms, "info", None)
        if info and "field" in info:
            field = info["field"]
        if field is None:
            return None
        f = np.asarray(field, dtype=float).reshape(-1)
        if f.size != 3:
            raise ValueError(f"field must have three components, got {f.size}")
        return (f * _BOHR / _HARTREE).tolist()  # V/A -> Hartree per e*Bohr

    def _resolve_state(self, atoms):
        p = self.parameters
        vals = {k: p[k] for k in ("charge", "multiplicity", "reference", "method")}
        info = getattr(atoms, "info", None) or {}
        for k in vals:
            if k in info:
                vals[k] = info[k]
        field = self._field_au(atoms)
        return (float(vals["charge"]), int(vals["multiplicity"]), str(vals["reference"]), str(vals["method"]),
                None if field is None else tuple(field))

    def check_state(self, atoms, tol=1e-15):
        changes = super().check_state(atoms, tol=tol)
        if self._resolve_state(atoms) != self._last_state and "info" not in changes:
            changes = list(changes) + ["info"]
        return changes

    def calculate(self, atoms=None, properties=("energy",), system_changes=all_changes):
        super().calculate(atoms, properties, system_changes)
        if np.any(self.atoms.pbc):
            raise NotImplementedError("periodic cells are not part of the am1g GPU engine")
        state = self._resolve_state(self.atoms)
        charge, multiplicity, reference, method, field = state
        numbers = self.atoms.get_atomic_numbers()
        positions = self.atoms.get_positions()
        if "forces" in properties:
            r = am1g.gradient(numbers, positions, charge, multiplicity, reference, method, field)
            self.results["forces"] = -np.asarray(r["gradient_ev_per_angstrom"], dtype=float)
        else:
            r = am1g.single_point(numbers, positions, charge, multiplicity, reference, method, field)
        self.results["energy"] = r["energy_ev"]
        self.results["free_energy"] = r["energy_ev"]
        self.results["charges"] = np.asarray(r["charges"], dtype=float)
        self.results["dipole"] = np.asarray(r["dipole_debye"], dtype=float) * Debye
        self.results["heat_of_formation_kcal"] = r.get("heat_of_formation_kcal")
        self._last_state = state
