"""The ASE calculator (am1g.ase.AM1) against the engine it wraps, with a minimal stand-in for the
three ASE names it imports (ASE is not installed in this image): forces are minus the gradient in
eV/A, energies eV, dipoles e*A, the field converted from V/A, atoms.info overrides the
constructor, periodic cells are refused."""
import sys
import types

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Atoms:
    def __init__(self, numbers, positions, pbc=False):
        self.numbers = np.array(numbers)
        self.positions = np.array(positions, dtype=float)
        self.pbc = np.array([pbc] * 3)
        self.info = {}

    def get_atomic_numbers(self):
        return self.numbers

    def get_positions(self):
        return self.positions

    def copy(self):
        a = _Atoms(self.numbers, self.positions)
        a.pbc = self.pbc.copy()
        a.info = dict(self.info)
        return a


@pytest.fixture
def ase_stub(monkeypatch):
    class Calculator:
        def __init__(self, **kwargs):
            self.parameters = dict(self.default_parameters)
            self.parameters.update(kwargs)
            self.results = {}
            self.atoms = None

        def check_state(self, atoms, tol=1e-15):
            if self.atoms is None or not np.array_equal(atoms.positions, self.atoms.positions):
                return ["positions"]
            return []

        def calculate(self, atoms=None, properties=("energy",), system_changes=()):
            if atoms is not None:
                self.atoms = atoms.copy()

        def get_property(self, name, atoms):
            if self.check_state(atoms) or name not in self.results:
                self.results = {}
                self.calculate(atoms, [name], [])
            return self.results[name]

    ase = types.ModuleType("ase")
    calc = types.ModuleType("ase.calculators")
    calc_calc = types.ModuleType("ase.calculators.calculator")
    calc_calc.Calculator = Calculator
    calc_calc.all_changes = ["positions", "numbers", "cell", "pbc", "info"]
    units = types.ModuleType("ase.units")
    units.Debye = 0.20819434
    for name, mod in (("ase", ase), ("ase.calculators", calc), ("ase.calculators.calculator", calc_calc),
                      ("ase.units", units)):
        monkeypatch.setitem(sys.modules, name, mod)
    monkeypatch.delitem(sys.modules, "am1g.ase", raising=False)
    import am1g.ase as adapter

    return adapter


WATER = ([8, 1, 1], [[0.0, 0.0, 0.1173], [0.0, 0.7572, -0.4692], [0.0, -0.7572, -0.4692]])


def test_energy_forces_dipole(ase_stub):
    import am1g

    atoms = _Atoms(*WATER)
    calc = ase_stub.AM1()
    f = calc.get_property("forces", atoms)
    g = am1g.gradient(*WATER)
    assert np.allclose(f, -np.array(g["gradient_ev_per_angstrom"]), atol=1e-12)
    assert calc.results["energy"] == g["energy_ev"]
    assert np.allclose(calc.results["dipole"], np.array(g["dipole_debye"]) * 0.20819434)
    sp = am1g.single_point(*WATER)
    assert abs(calc.get_property("energy", atoms) - sp["energy_ev"]) < 1e-9


def test_info_overrides_and_field(ase_stub):
    import am1g

    atoms = _Atoms(*WATER)
    atoms.info["charge"] = 1
    atoms.info["multiplicity"] = 2
    calc = ase_stub.AM1(field=[0.1, 0.0, 0.0])  # V/A
    e = calc.get_property("energy", atoms)
    ref = am1g.single_point(*WATER, 1.0, 2, electric_field=[0.1 * 0.529167 / 27.21, 0.0, 0.0])
    assert abs(e - ref["energy_ev"]) < 1e-9 and ref["unrestricted"]


def test_periodic_is_refused(ase_stub):
    atoms = _Atoms(*WATER, pbc=True)
    with pytest.raises(NotImplementedError):
        ase_stub.AM1().get_property("energy", atoms)
