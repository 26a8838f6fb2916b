"""The module-level ``orbitals`` and ``divide_conquer`` front ends (src/python.rs:1905-1992 and
448-663) and the fixed-density gradient entry point (am1g_gradient_density) on the GPU, against
the oracle on the same inputs."""
import numpy as np
import pytest

import am1g
from oracle import oracle as O
from tests import molecules as M

pytestmark = pytest.mark.gpu

ANG = 1.0 / 0.529167


def test_gradient_density_at_the_scf_density_is_the_scf_gradient(engine):
    for name, kw in (("FORMALDEHYDE", {}), ("METHYL", {"multiplicity": 2})):
        z, x = M.parse(getattr(M, name))
        r = engine.gradient(z, x, **kw)
        s = engine.scf(z, x, full=True, **kw)
        g = engine.gradient_density(z, x, s["density"], s.get("spin_density"), **kw)
        assert np.abs(g - r["gradient"]).max() < 1e-9, name
        o = O.gradient_density(z, x, s["density"], s.get("spin_density"), **kw)
        assert np.abs(g - o).max() < 1e-5, name


def test_orbitals_match_the_oracle():
    z, x = M.parse(M.FORMALDEHYDE)
    pos = x / ANG
    d = am1g.orbitals(z, pos)
    o = O.scf(z, x, full=True)
    assert np.abs(np.asarray(d["energies_ev"]) - o["mo_energies"]).max() < 1e-6
    assert np.allclose(np.asarray(d["energies_hartree"]) * 27.21, d["energies_ev"])
    c = np.asarray(d["coefficients"])
    occ = c[:, : d["n_occupied"]]
    assert np.abs(2.0 * occ @ occ.T - o["density"]).max() < 1e-6  # columns are the orbitals
    assert len(d["ao_labels"]) == c.shape[0] and d["ao_labels"][:2] == [(0, "s"), (0, "px")]
    assert not d["unrestricted"] and "beta_energies_ev" not in d
    u = am1g.orbitals(*M.parse(M.METHYL)[0:1], M.parse(M.METHYL)[1] / ANG, multiplicity=2)
    assert u["unrestricted"] and u["beta_n_occupied"] == u["n_occupied"] - 1


@pytest.mark.parametrize("multiplicity", [1, 2])
def test_divide_conquer_front_end(multiplicity):
    z, x = M.water_chain(6)
    if multiplicity == 2:
        keep = np.ones(len(z), dtype=bool)
        keep[4] = False
        z, x = z[keep], x[keep]
    kw = dict(core_size=3, buffer_radius=8.0, smearing_ev=0.05, e_tol=1e-8, p_tol=1e-7, max_scf=300)
    d = am1g.divide_conquer(z, x / ANG, multiplicity=multiplicity, **kw)
    o = O.dc(z, x, multiplicity=multiplicity, core_size=3, buffer_radius=8.0, kt=0.05, e_tol=1e-8, p_tol=1e-7,
             max_scf=300)
    assert abs(d["energy_ev"] - o["total_ev"]) <= 1e-6
    assert d["unrestricted"] == (multiplicity == 2)
    assert len(d["fermi_energies_ev"]) == (2 if multiplicity == 2 else 1)
    nat = len(z)
    nao = sum(1 if v == 1 else 4 for v in z)
    assert d["coulomb_work"] == nat * (nat - 1) // 2
    assert d["dense_triangle_elements"] == nao * (nao + 1) // 2
    assert 0 < d["exchange_work"] <= d["coulomb_work"] and d["retained_density_blocks"] <= nat * nat
    # forces: the fixed-density gradient at the GPU's own D&C density, against the oracle's
    eng = am1g.engine()
    r = eng.dc(z, x, multiplicity=multiplicity, core_size=3, buffer_radius=8.0, kt=0.05, e_tol=1e-8, p_tol=1e-7,
               max_scf=300, density=True)
    go = O.gradient_density(z, x, r["density"], r["spin_density"], multiplicity=multiplicity)
    f = np.asarray(d["forces_ev_per_angstrom"])
    assert np.abs(-f / ANG - go).max() < 1e-5
