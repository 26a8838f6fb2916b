"""The batched kernel's eigensolver for orders 129-256 (the tridiagonal method with the matrix in
global memory, wg_eigh_tri_g) against the oracle: water clusters of 132-252 basis functions,
a doublet radical (UHF) and RM1 in one launch, with energies, charges and frontier orbitals
within the 1e-6 contract, and the same molecules through the Jacobi solver (tri_max 128) for
comparison."""
import numpy as np
import pytest

from am1g import geometry as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(41)
    out = []
    for k in (22, 27, 33, 38, 42):
        m = G.small_water_cluster(k, rng)
        out.append((m.z, m.xyz_bohr, {}))
    m = G.small_water_cluster(30, rng)
    keep = np.ones(len(m.z), dtype=bool)
    keep[np.nonzero(m.z == 1)[0][3]] = False
    out.append((m.z[keep], m.xyz_bohr[keep], {"multiplicity": 2}))
    m = G.small_water_cluster(25, rng)
    out.append((m.z, m.xyz_bohr, {"method": "rm1"}))
    return out


def _check(r, o):
    assert r["status"] == 0 and o["status"] == 0, r.get("message")
    assert abs(r["total_ev"] - o["total_ev"]) <= 1e-6
    assert np.abs(r["charges"] - o["charges"]).max() <= 1e-6
    for k in ("homo_ev", "lumo_ev", "homo_beta_ev", "lumo_beta_ev"):
        if o[k] is not None:
            assert abs(r[k] - o[k]) <= 1e-6, k
    assert abs(r["iterations"] - o["iterations"]) <= 1


def test_batched_orders_129_to_256_match_the_oracle(engine):
    cases = _cases()
    old = engine.tunables()["tri_max"]
    try:
        engine.set_tunable("tri_max", 256)
        res = engine.scf_batch(cases, full=True)
        engine.set_tunable("tri_max", 128)
        jac = engine.scf_batch(cases)
    finally:
        engine.set_tunable("tri_max", old)
    for (z, x, kw), r, j in zip(cases, res, jac):
        o = O.scf(z, x, full=True, **kw)
        _check(r, o)
        _check(j, o)
        # orbitals: the density is the projector the eigenvectors define
        assert np.abs(r["density"] - o["density"]).max() <= 1e-6


def test_dc_subsystems_above_128(engine):
    """D&C subsystems of more than 128 basis functions take the same solver."""
    m = G.water_cluster((4, 4, 4), seed=3)
    kw = dict(core_size=12, buffer_radius=8.0, kt=0.05, e_tol=1e-8, p_tol=1e-7, max_scf=300)
    r = engine.dc(m.z, m.xyz_bohr, **kw)
    o = O.dc(m.z, m.xyz_bohr, **kw)
    assert r["status"] == 0 and o["status"] == 0
    assert r["largest_subsystem_aos"] == o["largest_subsystem_aos"] > 128
    assert abs(r["total_ev"] - o["total_ev"]) <= 1e-6
    assert np.abs(np.asarray(r["charges"]) - o["charges"]).max() <= 1e-6
