"""Nuclear gradient. The oracle's gradient (central differences of the energy at the converged
densities, oracle_gradient) is pinned here against what the mathematics fixes: central
differences of the fully converged SCF energy itself (the reference's own check,
tests/element_coverage.rs:101-162 and gradient.rs:95-128 numerical_gradient), translational
and rotational invariance, step-size independence and the field force -F q (tests/external_field.rs:
167-190). The GPU gradient (am1g_gradient, through the C ABI) is compared with the oracle's."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import molecules as M

TIGHT = dict(e_tol=1e-12, p_tol=1e-12)


def _scf_fd(z, x, h=1e-4, **kw):
    g = np.zeros_like(x)
    for a in range(len(z)):
        for k in range(3):
            xp = x.copy()
            xm = x.copy()
            xp[a, k] += h
            xm[a, k] -= h
            ep = O.scf(z, xp, full=False, **TIGHT, **kw)["total_ev"]
            em = O.scf(z, xm, full=False, **TIGHT, **kw)["total_ev"]
            g[a, k] = (ep - em) / (2 * h)
    return g


@pytest.mark.parametrize("name,kw", [("water", {}), ("formaldehyde", {}), ("methyl", {"multiplicity": 2}),
                                     ("water_rm1", {"method": "rm1"})])
def test_oracle_gradient_is_the_derivative_of_the_scf_energy(name, kw):
    src = {"water": M.WATER, "formaldehyde": M.FORMALDEHYDE, "methyl": M.METHYL, "water_rm1": M.WATER}[name]
    z, x = M.parse(src)
    r = O.gradient(z, x, **TIGHT, **kw)
    assert r["status"] == 0
    g = r["gradient"]
    assert np.abs(g - _scf_fd(z, x, **kw)).max() < 5e-6  # eV/Bohr
    assert np.abs(g.sum(axis=0)).max() < 1e-8  # no net force
    torque = np.cross(x, g).sum(axis=0)
    assert np.abs(torque).max() < 1e-7  # no net torque
    g2 = O.gradient(z, x, step=5e-5, **TIGHT, **kw)["gradient"]
    assert np.abs(g - g2).max() < 2e-9  # five-point differences: no step dependence left


def test_field_force_is_minus_charge_times_field():
    # tests/external_field.rs:167-190: with a field the net force is Q F (here Q = +1)
    z, x = M.parse(M.WATER)
    f = np.array([0.05, -0.02, 0.03])
    r = O.gradient(z, x, charge=1.0, multiplicity=2, electric_field=f, **TIGHT)
    assert r["status"] == 0
    assert np.allclose(r["gradient"].sum(axis=0), -f * 1.0, atol=1e-7)
    assert np.abs(r["gradient"] - _scf_fd(z, x, charge=1.0, multiplicity=2, electric_field=f)).max() < 5e-6


def test_planted_sign_error_is_caught():
    # a gradient with one component's sign flipped fails the finite-difference check
    z, x = M.parse(M.FORMALDEHYDE)
    g = O.gradient(z, x, **TIGHT)["gradient"].copy()
    g[1, 2] = -g[1, 2]
    assert np.abs(g - _scf_fd(z, x)).max() > 1e-3


GPU_TOL = 1e-6  # eV/Bohr against the oracle's gradient


@pytest.mark.gpu
@pytest.mark.parametrize("name,kw", [("water", {}), ("formaldehyde", {}), ("methyl", {"multiplicity": 2}),
                                     ("ch3cl", {}), ("water_rm1", {"method": "rm1"}),
                                     ("water_field", {"electric_field": [0.05, -0.02, 0.03]})])
def test_gpu_gradient_matches_oracle(engine, name, kw):
    src = {"water": M.WATER, "formaldehyde": M.FORMALDEHYDE, "methyl": M.METHYL, "water_rm1": M.WATER,
           "ch3cl": M.RECORDED_HOF["CH3Cl"][0], "water_field": M.WATER}[name]
    z, x = M.parse(src)
    o = O.gradient(z, x, **kw)
    r = engine.gradient(z, x, **kw)
    assert r["status"] == 0 and o["status"] == 0
    assert abs(r["total_ev"] - o["total_ev"]) <= 1e-6
    assert np.abs(r["gradient"] - o["gradient"]).max() <= GPU_TOL


@pytest.mark.gpu
def test_gpu_gradient_chain_and_large_path(engine, monkeypatch):
    """A chain on the one-workgroup SCF and a 40-water cluster (240 basis functions) on the
    grid-level SCF; the gradient kernels are the same."""
    from am1g import geometry as G

    m = G.chain_batch(4, seed=7)[3]
    o = O.gradient(m.z, m.xyz_bohr)
    r = engine.gradient(m.z, m.xyz_bohr)
    assert np.abs(r["gradient"] - o["gradient"]).max() <= GPU_TOL
    w = G.small_water_cluster(40, np.random.default_rng(3))
    o = O.gradient(w.z, w.xyz_bohr)
    r = engine.gradient(w.z, w.xyz_bohr)
    assert r["nao"] == 240 and np.abs(r["gradient"] - o["gradient"]).max() <= GPU_TOL
    assert np.abs(r["gradient"].sum(axis=0)).max() < 1e-7


@pytest.mark.gpu
def test_python_gradient_contract():
    import am1g

    pos = [[0.0, 0.0, 0.1173], [0.0, 0.7572, -0.4692], [0.0, -0.7572, -0.4692]]
    r = am1g.gradient([8, 1, 1], pos)
    keys = {"energy_hartree", "energy_ev", "heat_of_formation_kcal", "electronic_ev", "core_ev", "homo_ev",
            "lumo_ev", "gradient_hartree_per_bohr", "gradient_ev_per_angstrom", "charges", "dipole_debye",
            "dipole_magnitude", "max_gradient_hartree_per_bohr", "converged", "iterations", "unrestricted", "method"}
    assert keys <= set(r)
    o = O.gradient([8, 1, 1], np.array(pos) * am1g.ANGSTROM_TO_BOHR)
    g_au = np.array(r["gradient_hartree_per_bohr"])
    assert np.abs(g_au * 27.21 - o["gradient"]).max() <= GPU_TOL
    assert np.allclose(np.array(r["gradient_ev_per_angstrom"]), o["gradient"] / 0.529167, atol=1e-5)
