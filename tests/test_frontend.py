"""Python front end: ``am1g.single_point`` keeps am1-rs's ``native.single_point`` contract
(python/am1_rs/native.py:52-97, src/python.rs:147-199): signature and positional order, the
result keys the reference's own contract test requires (tests/test_result_contract.py:42-55),
``ValueError`` for every bad input (python.rs:27-29)."""
import inspect

import numpy as np
import pytest

import am1g
from oracle import oracle as O

# tests/test_result_contract.py:42-55: the molecular energy report
ENERGY_KEYS = {"energy_ev", "energy_hartree", "electronic_ev", "core_ev", "heat_of_formation_kcal", "homo_ev",
               "lumo_ev", "dipole_debye", "dipole_magnitude", "charges", "iterations", "unrestricted"}
# python.rs:166-198: every key single_point sets
ALL_KEYS = ENERGY_KEYS | {"method", "homo_beta_ev", "lumo_beta_ev", "field_nuclear_ev", "converged"}

WATER_NUMBERS = [8, 1, 1]
WATER_POSITIONS = [[0.0, 0.0, 0.1173], [0.0, 0.7572, -0.4692], [0.0, -0.7572, -0.4692]]


def test_signature_matches_the_reference():
    params = list(inspect.signature(am1g.single_point).parameters.values())
    assert [p.name for p in params] == ["numbers", "positions", "charge", "multiplicity", "reference", "method",
                                        "electric_field"]
    assert [p.default for p in params[2:]] == [0.0, 1, "auto", "am1", None]


@pytest.mark.parametrize("numbers", [[262, 1, 1], [0, 1, 1], [8.5, 1, 1], [-1, 1, 1]])
def test_out_of_range_atomic_numbers_raise(numbers):
    # 262 would wrap to 6 (carbon) in a plain uint8 cast
    with pytest.raises(ValueError):
        am1g.single_point(numbers, WATER_POSITIONS)


def test_shape_errors_raise():
    with pytest.raises(ValueError):
        am1g.single_point([8, 1], WATER_POSITIONS)
    with pytest.raises(ValueError):
        am1g.single_point(WATER_NUMBERS, [[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    with pytest.raises(ValueError):
        am1g.single_point(WATER_NUMBERS, WATER_POSITIONS, electric_field=[0.0, 0.0])
    with pytest.raises(ValueError):
        am1g.single_point(WATER_NUMBERS, WATER_POSITIONS, method="pm3")


def test_engine_request_rejects_wrapping_numbers():
    eng = am1g.Engine.__new__(am1g.Engine)
    eng._L = am1g.load()
    with pytest.raises(ValueError):
        eng._request(np.array([262, 1, 1]), np.zeros((3, 3)), [])


@pytest.mark.gpu
def test_single_point_keys_and_values():
    r = am1g.single_point(WATER_NUMBERS, WATER_POSITIONS)
    assert ENERGY_KEYS <= set(r)
    assert ALL_KEYS <= set(r)
    xyz = np.array(WATER_POSITIONS) * am1g.ANGSTROM_TO_BOHR
    e = am1g.engine().scf(WATER_NUMBERS, xyz, full=False)
    assert r["energy_ev"] == e["total_ev"]
    assert r["energy_hartree"] == pytest.approx(e["total_ev"] / 27.21, rel=1e-15)
    assert r["electronic_ev"] == e["electronic_ev"] and r["core_ev"] == e["core_ev"]
    assert r["charges"] == e["charges"].tolist()
    assert r["dipole_magnitude"] == pytest.approx(float(np.linalg.norm(r["dipole_debye"])), rel=1e-14)
    assert r["method"] == "am1" and r["converged"] and not r["unrestricted"]
    assert r["homo_beta_ev"] is None and r["lumo_beta_ev"] is None and r["field_nuclear_ev"] == 0.0
    o = O.scf(WATER_NUMBERS, xyz, full=False)
    assert abs(r["energy_ev"] - o["total_ev"]) <= 1e-6


@pytest.mark.gpu
def test_single_point_positional_reference_and_method():
    # positional order (numbers, positions, charge, multiplicity, reference, method)
    methyl = [[0.0, 0.0, 0.0], [1.079, 0.0, 0.0], [-0.5395, 0.9344, 0.0], [-0.5395, -0.9344, 0.0]]
    r = am1g.single_point([6, 1, 1, 1], methyl, 0.0, 2, "uhf", "rm1")
    assert r["unrestricted"] and r["method"] == "rm1"
    assert r["homo_beta_ev"] is not None
    o = O.scf([6, 1, 1, 1], np.array(methyl) * am1g.ANGSTROM_TO_BOHR, full=False, multiplicity=2, reference="uhf",
              method="rm1")
    assert abs(r["energy_ev"] - o["total_ev"]) <= 1e-6
    assert abs(r["homo_beta_ev"] - o["homo_beta_ev"]) <= 1e-6


@pytest.mark.gpu
def test_single_point_electric_field():
    """A field in atomic units (python.rs:55-70) through single_point against the oracle with the
    same field in eV per (e*Bohr)."""
    f_au = np.array([0.002, -0.001, 0.003])
    r = am1g.single_point(WATER_NUMBERS, WATER_POSITIONS, electric_field=f_au)
    xyz = np.array(WATER_POSITIONS) * am1g.ANGSTROM_TO_BOHR
    o = O.scf(WATER_NUMBERS, xyz, full=False, electric_field=f_au * 27.21)
    assert abs(r["energy_ev"] - o["total_ev"]) <= 1e-6
    assert abs(r["field_nuclear_ev"] - o["field_nuclear_ev"]) <= 1e-12
    assert np.abs(np.array(r["charges"]) - o["charges"]).max() <= 1e-6
    assert abs(r["homo_ev"] - o["homo_ev"]) <= 1e-6
    z0 = am1g.single_point(WATER_NUMBERS, WATER_POSITIONS)
    assert abs(z0["energy_ev"] - r["energy_ev"]) > 1e-3  # the field does something


@pytest.mark.gpu
def test_electric_field_large_path_and_batch(engine, monkeypatch):
    """The field in the batched kernel (beside a field-free copy of the molecule in the same
    launch) and on the grid-level path, against the oracle."""
    from tests import molecules as M

    z, x = M.parse(M.FORMALDEHYDE)
    f = np.array([0.05, 0.02, -0.04])
    o = O.scf(z, x, full=False, electric_field=f)
    rb = engine.scf_batch([(z, x, {"electric_field": f}), (z, x)])
    assert abs(rb[0]["total_ev"] - o["total_ev"]) <= 1e-6
    assert np.abs(rb[0]["charges"] - o["charges"]).max() <= 1e-6
    assert rb[0]["field_nuclear_ev"] == pytest.approx(o["field_nuclear_ev"], abs=1e-12)
    assert rb[1]["field_nuclear_ev"] == 0.0
    monkeypatch.setenv("AM1G_FORCE_LARGE_PATH", "1")
    rl = engine.scf(z, x, full=False, electric_field=f)
    assert abs(rl["total_ev"] - o["total_ev"]) <= 1e-6
    assert np.abs(rl["charges"] - o["charges"]).max() <= 1e-6


def test_orbitals_and_divide_conquer_signatures_match_the_reference():
    # python.rs:1910 (orbitals) and :448-460 (divide_conquer)
    p = list(inspect.signature(am1g.orbitals).parameters.values())
    assert [x.name for x in p] == ["numbers", "positions", "charge", "multiplicity", "reference", "method",
                                   "electric_field"]
    p = list(inspect.signature(am1g.divide_conquer).parameters.values())
    assert [x.name for x in p] == ["numbers", "positions", "charge", "multiplicity", "reference", "method",
                                   "core_size", "buffer_radius", "smearing_ev", "e_tol", "p_tol", "max_scf", "mixing",
                                   "gap_warn_ev", "forces", "multipole_cutoff", "electric_field", "cell", "pbc",
                                   "realspace_cutoff", "exchange_cutoff", "optimize", "opt_max_iter", "opt_gtol"]
    assert [x.default for x in p[2:]] == [0.0, 1, "auto", "am1", 12, 11.0, 0.05, 1.0e-7, 1.0e-6, 300, 0.4, 0.5, True,
                                          None, None, None, None, 40.0, 20.0, False, 200, 1.0e-3]


def test_divide_conquer_refuses_cells_and_optimization():
    with pytest.raises(ValueError):  # cell without pbc is ambiguous (python.rs:494-500)
        am1g.divide_conquer(WATER_NUMBERS, WATER_POSITIONS, cell=[[10, 0, 0], [0, 10, 0], [0, 0, 10]])
    with pytest.raises(NotImplementedError):
        am1g.divide_conquer(WATER_NUMBERS, WATER_POSITIONS, cell=[[10, 0, 0], [0, 10, 0], [0, 0, 10]],
                            pbc=[True, True, True])
    with pytest.raises(NotImplementedError):
        am1g.divide_conquer(WATER_NUMBERS, WATER_POSITIONS, optimize=True)


def test_oracle_fixed_density_gradient_is_the_scf_gradient_at_its_density():
    """At the converged SCF density the fixed-density gradient is the SCF gradient (the two
    oracle routines differentiate the same energy at the same densities)."""
    from tests import molecules as M

    for name, kw in (("formaldehyde", {}), ("methyl", {"multiplicity": 2})):
        z, x = M.parse(getattr(M, name.upper()))
        r = O.scf(z, x, full=True, e_tol=1e-11, p_tol=1e-10, **kw)
        g = O.gradient(z, x, e_tol=1e-11, p_tol=1e-10, **kw)["gradient"]
        sp = r.get("spin_density")
        gd = O.gradient_density(z, x, r["density"], sp, **kw)
        assert np.abs(gd - g).max() < 1e-9, name


def test_oracle_dc_full_buffer_forces_are_the_scf_forces():
    """divide_conquer_gradient at a full-buffer D&C density (which equals the SCF density,
    tests/divide_conquer.rs:160-214) is the SCF gradient."""
    from tests import molecules as M

    z, x = M.water_chain(3)
    dc = O.dc(z, x, core_size=3, buffer_radius=500.0, kt=0.0, e_tol=1e-11, p_tol=1e-10, max_scf=400, density=True)
    g_dc = O.gradient_density(z, x, dc["density"])
    g = O.gradient(z, x, e_tol=1e-11, p_tol=1e-10)["gradient"]
    assert np.abs(g_dc - g).max() < 1e-6
