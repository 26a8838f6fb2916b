"""GPU divide-and-conquer SCF (``am1g_dc``) against the oracle's divide-and-conquer on the same
inputs: the same partition, Fermi level, Yang assembly and density DIIS, so energies, charges and
the Fermi level agree to the engine's tolerance (1e-6 eV, 1e-6 e) and iteration counts within one.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import molecules as M

pytestmark = pytest.mark.gpu

E_TOL = 1e-6
Q_TOL = 1e-6
KT_1000K = 0.08617333262  # eV: k_B * 1000 K


def _compare(g, o, iter_slack=1):
    assert g["status"] == o["status"] == 0, (g.get("message"), o.get("message"))
    assert g["n_subsystems"] == o["n_subsystems"]
    assert g["largest_subsystem_aos"] == o["largest_subsystem_aos"]
    assert abs(g["total_ev"] - o["total_ev"]) <= E_TOL
    assert abs(g["heat_of_formation_kcal"] - o["heat_of_formation_kcal"]) <= E_TOL * 23.061
    assert np.abs(g["charges"] - o["charges"]).max() <= Q_TOL
    assert abs(g["fermi_ev"] - o["fermi_ev"]) <= 1e-5
    assert abs(g["entropy_ev"] - o["entropy_ev"]) <= E_TOL
    assert abs(g["iterations"] - o["iterations"]) <= iter_slack


def test_water_chain_bisection_cores(engine):
    z, x = M.water_chain(8)
    kw = dict(core_size=6, buffer_radius=9.0, kt=0.05)
    _compare(engine.dc(z, x, **kw), O.dc(z, x, **kw))


def test_water_chain_molecule_cores_aufbau(engine):
    z, x = M.water_chain(6)
    kw = dict(core_of_atom=np.repeat(np.arange(6), 3), buffer_radius=7.5591, kt=0.0)
    _compare(engine.dc(z, x, **kw), O.dc(z, x, **kw))


def test_water_cluster_32_one_molecule_cores(engine):
    """The 2000-water workload's options (one-molecule cores, 4.0 A buffer, 1000 K) on the
    32-water cluster, where the oracle finishes in seconds."""
    from am1g import geometry as G

    mol = G.water_cluster_config(0)
    kw = dict(core_of_atom=np.arange(len(mol.z)) // 3, buffer_radius=4.0 / 0.529167, kt=KT_1000K)
    g = engine.dc(mol.z, mol.xyz_bohr, density=True, **kw)
    o = O.dc(mol.z, mol.xyz_bohr, density=True, **kw)
    _compare(g, o)
    assert np.abs(g["density"] - o["density"]).max() < 1e-6
    assert np.array_equal(g["density"], g["density"].T)


def test_full_buffer_reproduces_plain_scf(engine):
    # tests/divide_conquer.rs:160-214: a buffer covering the system gives the full SCF
    z, x = M.water_chain(4)
    full = engine.scf(z, x, e_tol=1e-11, p_tol=1e-10, full=False)
    dc = engine.dc(z, x, core_size=3, buffer_radius=500.0, kt=0.0, e_tol=1e-11, p_tol=1e-10, max_scf=400)
    assert dc["converged"]
    assert abs(dc["total_ev"] - full["total_ev"]) < 1e-6
    assert np.abs(dc["charges"] - full["charges"]).max() < 1e-6


def test_repeat_call_reuses_buffers_and_is_deterministic(engine):
    z, x = M.water_chain(5)
    a = engine.dc(z, x, core_size=6, buffer_radius=9.0)
    b = engine.dc(*M.water_chain(3), core_size=3, buffer_radius=6.0)
    c = engine.dc(z, x, core_size=6, buffer_radius=9.0)
    assert b["status"] == 0
    assert a["total_ev"] == c["total_ev"] and np.array_equal(a["charges"], c["charges"])


def test_error_paths(engine):
    z, x = M.water_chain(3)
    assert engine.dc(z, x, buffer_radius=0.0)["status"] == 1
    assert engine.dc(z, x, multiplicity=2)["status"] == 1  # even electron count, doublet
    assert engine.dc(np.array([2, 1, 1]), x[:3])["status"] == 2


def _radical_chain(n, drop=4):
    z, x = M.water_chain(n)
    keep = np.ones(len(z), dtype=bool)
    keep[drop] = False  # one hydrogen less: a neutral doublet
    return z[keep], x[keep]


def test_open_shell_full_buffer_matches_uhf(engine):
    # tests/divide_conquer.rs:305-353: two chemical potentials reproduce the UHF doublet
    z, x = _radical_chain(3)
    full = engine.scf(z, x, multiplicity=2, e_tol=1e-11, p_tol=1e-10, full=False)
    kw = dict(multiplicity=2, core_size=3, buffer_radius=500.0, kt=0.0, e_tol=1e-11, p_tol=1e-10, max_scf=400)
    g = engine.dc(z, x, **kw)
    assert g["unrestricted"] and g["converged"]
    assert abs(g["total_ev"] - full["total_ev"]) < 1e-6
    _compare(g, O.dc(z, x, **kw))


@pytest.mark.parametrize("kt", [0.0, 0.05])
def test_open_shell_partial_buffer(engine, kt):
    z, x = _radical_chain(6)
    kw = dict(multiplicity=2, core_size=6, buffer_radius=9.0, kt=kt)
    g = engine.dc(z, x, **kw)
    o = O.dc(z, x, **kw)
    assert g["unrestricted"]
    _compare(g, o)
    assert abs(g["fermi_beta_ev"] - o["fermi_beta_ev"]) <= 1e-5
    assert abs(g["charges"].sum()) < 1e-8


def test_water2000_full_size_against_oracle_fixture(engine):
    """The 2000-water workload at full size (6000 atoms, 12,000 AOs, 2000 one-molecule cores,
    4.0 A buffer, 1000 K) against the oracle's result committed in tests/golden (made by
    tests/golden/make_dc_water2000.py with the same tight tolerances)."""
    import os

    from am1g import geometry as G

    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "dc_water2000.npz"))
    mol = G.water_cluster_config(3)
    r = engine.dc(mol.z, mol.xyz_bohr, core_of_atom=np.arange(len(mol.z)) // 3, buffer_radius=4.0 / 0.529167,
                  kt=KT_1000K, e_tol=1e-9, p_tol=1e-10, max_scf=300)
    assert r["status"] == 0, r.get("message")
    assert r["n_subsystems"] == int(g["n_subsystems"]) == 2000
    assert r["largest_subsystem_aos"] == int(g["largest_subsystem_aos"])
    assert abs(r["total_ev"] - float(g["total_ev"])) <= E_TOL
    assert abs(r["heat_of_formation_kcal"] - float(g["heat_of_formation_kcal"])) <= E_TOL * 23.061
    assert np.abs(r["charges"] - g["charges"]).max() <= Q_TOL
    assert abs(r["charges"].sum()) < 1e-8
    assert abs(r["fermi_ev"] - float(g["fermi_ev"])) <= 1e-5


def test_external_field(engine):
    """The field enters the D&C through H_core and its nuclear term (divide_conquer.rs:494-500)."""
    z, x = M.water_chain(6)
    kw = dict(core_size=3, buffer_radius=8.0, kt=0.05, e_tol=1e-8, p_tol=1e-7, max_scf=300,
              electric_field=[0.02, -0.01, 0.03])
    r = engine.dc(z, x, **kw)
    o = O.dc(z, x, **kw)
    assert r["status"] == 0 and o["status"] == 0
    assert abs(r["total_ev"] - o["total_ev"]) <= 1e-6
    assert np.abs(np.asarray(r["charges"]) - o["charges"]).max() <= 1e-6
