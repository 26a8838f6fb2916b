"""The large-system SCF path (one molecule on the whole device: grid Fock kernels, library GEMMs,
device DIIS histories, the large-order eigensolver) against the FP64 oracle. Molecules above 256
atoms or basis functions take it automatically; AM1G_FORCE_LARGE_PATH=1 sends small ones there
too, so it is checked where the oracle is quick."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import molecules as M
from tests.test_gpu_parity import _compare

pytestmark = pytest.mark.gpu


@pytest.fixture
def forced(monkeypatch):
    monkeypatch.setenv("AM1G_FORCE_LARGE_PATH", "1")


@pytest.mark.parametrize("acc", ["none", "cdiis", "adiis_cdiis"])
def test_formaldehyde_accelerators(engine, forced, acc):
    z, x = M.parse(M.FORMALDEHYDE)
    r = engine.scf(z, x, accelerator=acc)
    _compare(r, O.scf(z, x, accelerator=acc))
    if acc == "adiis_cdiis":
        assert r["iterations"] == 11  # scf.rs:1039-1040


@pytest.mark.parametrize("name", ["H2O", "CH3Cl", "HCN"])
def test_recorded_molecules(engine, forced, name):
    z, x = M.parse(M.RECORDED_HOF[name][0])
    r = engine.scf(z, x)
    _compare(r, O.scf(z, x))
    assert abs(r["heat_of_formation_kcal"] - M.RECORDED_HOF[name][1]) < 1e-4


def test_density_and_orbitals(engine, forced):
    z, x = M.parse(M.FORMALDEHYDE)
    r = engine.scf(z, x)
    o = O.scf(z, x)
    assert np.abs(r["density"] - o["density"]).max() < 1e-6
    assert np.abs(r["mo_energies"] - o["mo_energies"]).max() < 1e-6
    C = r["mo_coeff"]
    assert np.allclose(C.T @ C, np.eye(C.shape[0]), atol=1e-10)


def test_water_cluster_32(engine, forced):
    from am1g import geometry as G

    mol = G.water_cluster_config(0)
    r = engine.scf(mol.z, mol.xyz_bohr, accelerator="cdiis")
    _compare(r, O.scf(mol.z, mol.xyz_bohr, accelerator="cdiis"))


def test_water_cluster_100_takes_the_large_path(engine):
    """300 atoms, 600 basis functions: beyond the batched kernel, so the large path runs without
    forcing. FP64 against the oracle with the same options; mixed precision (another trajectory,
    ended by a density-change test 100x tighter than p_tol) against the oracle run with that same
    tighter test, both to the full contract."""
    from am1g import geometry as G

    mol = G.water_cluster((5, 5, 4), seed=4)
    o = O.scf(mol.z, mol.xyz_bohr, full=False)
    r = engine.scf(mol.z, mol.xyz_bohr, full=False)
    _compare(r, o)
    rm = engine.scf(mol.z, mol.xyz_bohr, full=False, precision="mixed")
    assert rm["status"] == 0 and rm["fp32_iterations"] > 0
    _compare(rm, O.scf(mol.z, mol.xyz_bohr, full=False, p_tol=1e-9), iter_slack=99)


def test_methyl_radical_uhf(engine, forced):
    # scf.rs:1372-1388
    z, x = M.parse(M.METHYL)
    r = engine.scf(z, x, multiplicity=2)
    o = O.scf(z, x, multiplicity=2)
    assert r["unrestricted"]
    _compare(r, o)
    assert np.abs(r["mo_energies_beta"] - o["mo_energies_beta"]).max() < 1e-6
    assert abs(np.trace(r["spin_density"]) - 1.0) < 1e-9


def test_forced_uhf_singlet_equals_rhf(engine, forced):
    # scf.rs:1391-1418
    z, x = M.parse(M.WATER)
    u = engine.scf(z, x, reference="uhf")
    r = engine.scf(z, x)
    assert u["unrestricted"] and abs(u["total_ev"] - r["total_ev"]) < 1e-6
    _compare(u, O.scf(z, x, reference="uhf"))


def test_water_cluster_radical_uhf_large(engine):
    """A 300-atom doublet (100 waters less one hydrogen): UHF on the large path, FP64 and mixed."""
    from am1g import geometry as G

    mol = G.water_cluster((5, 5, 4), seed=4)
    keep = np.ones(len(mol.z), dtype=bool)
    keep[2] = False
    z, x = mol.z[keep], mol.xyz_bohr[keep]
    o = O.scf(z, x, multiplicity=2, full=False)
    r = engine.scf(z, x, multiplicity=2, full=False)
    assert r["unrestricted"]
    _compare(r, o)
    rm = engine.scf(z, x, multiplicity=2, full=False, precision="mixed")
    assert rm["status"] == 0 and abs(rm["total_ev"] - o["total_ev"]) <= 1e-6
    _compare(rm, O.scf(z, x, multiplicity=2, full=False, p_tol=1e-9), iter_slack=99)


def test_water512_full_size_against_oracle_fixture(engine):
    """The 512-water workload at full size (1536 atoms, 3072 basis functions) against the
    oracle's result committed in tests/golden (tests/golden/make_scf_water512.py): FP64 to the
    parity contract; mixed precision to the same contract (1e-6 eV in energy and in the frontier
    orbital energies, 1e-6 e in charges) against the oracle run with the mixed run's tighter
    density test (p_tol 1e-9, tests/golden/scf_water512_tight.npz)."""
    import os

    from am1g import geometry as G

    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "scf_water512.npz"))
    mol = G.water_cluster_config(2)
    r = engine.scf(mol.z, mol.xyz_bohr, full=False)
    assert r["status"] == 0, r.get("message")
    assert abs(r["total_ev"] - float(g["total_ev"])) <= 1e-6
    assert abs(r["heat_of_formation_kcal"] - float(g["heat_of_formation_kcal"])) <= 1e-6 * 23.061
    assert np.abs(r["charges"] - g["charges"]).max() <= 1e-6
    assert abs(r["homo_ev"] - float(g["homo_ev"])) <= 1e-6 and abs(r["lumo_ev"] - float(g["lumo_ev"])) <= 1e-6
    assert abs(r["iterations"] - int(g["iterations"])) <= 1
    assert np.allclose(r["dipole_debye"], g["dipole_debye"], atol=1e-5)
    rm = engine.scf(mol.z, mol.xyz_bohr, full=False, precision="mixed")
    assert rm["status"] == 0 and rm["fp32_iterations"] > 0
    gf = np.load(os.path.join(os.path.dirname(__file__), "golden", "scf_water512_tight.npz"))
    assert abs(rm["total_ev"] - float(gf["total_ev"])) <= 1e-6
    assert np.abs(rm["charges"] - gf["charges"]).max() <= 1e-6
    assert abs(rm["homo_ev"] - float(gf["homo_ev"])) <= 1e-6
    assert abs(rm["lumo_ev"] - float(gf["lumo_ev"])) <= 1e-6


def test_sp2_densities_match_diagonalization(engine):
    """The grid-level SCF forms each iteration's density by SP2 purification of the extrapolated
    Fock matrix (tunable sp2, on by default; eigen_sweeps counts the purification steps) and
    diagonalizes once, after the loop. The same SCF diagonalizing in every iteration reaches the
    same fixed point: 72 waters (432 basis functions), RHF and a UHF doublet (one more hydrogen
    far away), energies, charges, orbital energies and densities."""
    from am1g import geometry as G

    mol = G.water_cluster((6, 6, 2), seed=1)
    radical = (np.concatenate([mol.z, [1]]), np.concatenate([mol.xyz_bohr, [[40.0, 40.0, 40.0]]]))
    for z, x, kw in ((mol.z, mol.xyz_bohr, {}), (radical[0], radical[1], {"multiplicity": 2})):
        try:
            engine.set_tunable("sp2", 0)
            a = engine.scf(z, x, **kw)
            engine.set_tunable("sp2", 1)
            b = engine.scf(z, x, **kw)
        finally:
            engine.set_tunable("sp2", 1)
        assert a["status"] == 0 and b["status"] == 0
        assert a["eigen_sweeps"] == 0 and b["eigen_sweeps"] > 0
        assert abs(a["iterations"] - b["iterations"]) <= 1
        assert abs(a["total_ev"] - b["total_ev"]) <= 1e-8
        assert np.abs(a["charges"] - b["charges"]).max() <= 1e-9
        assert abs(a["homo_ev"] - b["homo_ev"]) <= 1e-8 and abs(a["lumo_ev"] - b["lumo_ev"]) <= 1e-8
        assert np.abs(a["mo_energies"] - b["mo_energies"]).max() <= 1e-8
        assert np.abs(a["density"] - b["density"]).max() <= 1e-9
        if kw:
            assert abs(a["homo_beta_ev"] - b["homo_beta_ev"]) <= 1e-8
            assert np.abs(a["mo_energies_beta"] - b["mo_energies_beta"]).max() <= 1e-8


def test_sp2_declines_a_degenerate_fermi_level(engine, forced):
    """Two waters 50 Angstrom apart with two electrons removed (a closed shell): the Fermi level
    falls inside the pair of degenerate HOMOs, so the purification cannot reach the electron
    count, and the iteration diagonalizes instead; after one iteration the result equals the
    diagonalizing run's bit for bit."""
    z, x = M.parse(M.WATER)
    z2 = np.concatenate([z, z])
    x2 = np.concatenate([x, x + np.array([50.0 * M.ANGSTROM_TO_BOHR, 0.0, 0.0])])
    out = []
    try:
        for sp2 in (0, 1):
            engine.set_tunable("sp2", sp2)
            out.append(engine.scf(z2, x2, charge=2.0, max_scf=1))
    finally:
        engine.set_tunable("sp2", 1)
    a, b = out
    assert a["iterations"] == b["iterations"] == 1
    assert a["total_ev"] == b["total_ev"]
    assert np.array_equal(a["charges"], b["charges"])
    assert np.array_equal(a["mo_energies"], b["mo_energies"])
    assert np.array_equal(a["density"], b["density"])


def test_frontier_orbital_energies_by_bisection(engine):
    """When the caller takes neither orbital energies nor orbitals, the grid-level SCF finds HOMO
    and LUMO by bisection on the tridiagonal form of its last Fock matrix instead of the full
    eigensolve; they equal the full eigensolve's (RHF and a UHF doublet, 72 waters)."""
    from am1g import geometry as G

    mol = G.water_cluster((6, 6, 2), seed=1)
    radical = (np.concatenate([mol.z, [1]]), np.concatenate([mol.xyz_bohr, [[40.0, 40.0, 40.0]]]))
    for z, x, kw in ((mol.z, mol.xyz_bohr, {}), (radical[0], radical[1], {"multiplicity": 2})):
        a = engine.scf(z, x, full=True, **kw)
        b = engine.scf(z, x, full=False, **kw)
        assert a["status"] == 0 and b["status"] == 0
        assert a["total_ev"] == b["total_ev"]
        keys = ("homo_ev", "lumo_ev") + (("homo_beta_ev", "lumo_beta_ev") if kw else ())
        for k in keys:
            assert abs(a[k] - b[k]) <= 1e-10, (k, a[k], b[k])
