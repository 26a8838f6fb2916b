"""Pins on the oracle's divide-and-conquer SCF (CPU only): the identities the method rests on,
as the reference's own tests state them (tests/divide_conquer.rs)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import molecules as M


@pytest.mark.parametrize("buffer", [6.0, 9.0, 14.0, 40.0])
def test_partition_weights_sum_to_one(buffer):
    # tests/divide_conquer.rs:94-137
    z, x = M.water_chain(9)
    W = O.dc_weight_sum(z, x, core_size=6, buffer_radius=buffer)
    d = np.linalg.norm(x[:, None, :] - x[None, :, :], axis=-1)
    kept = d <= buffer
    assert np.abs(W[kept] - 1.0).max() < 1e-12
    assert np.all(W[~kept] == 0.0)


def test_molecule_cores_also_sum_to_one():
    z, x = M.water_chain(6)
    cores = np.repeat(np.arange(6), 3)
    W = O.dc_weight_sum(z, x, buffer_radius=7.5591, core_of_atom=cores)
    d = np.linalg.norm(x[:, None, :] - x[None, :, :], axis=-1)
    assert np.abs(W[d <= 7.5591] - 1.0).max() < 1e-12
    assert np.all(W[d > 7.5591] == 0.0)


def test_full_buffer_reproduces_full_scf():
    # tests/divide_conquer.rs:160-214 (aufbau filling, buffer covering the system)
    z, x = M.water_chain(4)
    full = O.scf(z, x, e_tol=1e-11, p_tol=1e-10, full=False)
    dc = O.dc(z, x, core_size=3, buffer_radius=500.0, kt=0.0, e_tol=1e-11, p_tol=1e-10, max_scf=400)
    assert dc["converged"]
    assert abs(dc["total_ev"] - full["total_ev"]) < 1e-7
    assert np.abs(dc["charges"] - full["charges"]).max() < 1e-7


def test_open_shell_full_buffer_matches_uhf():
    # tests/divide_conquer.rs:305-353: two chemical potentials reproduce the UHF doublet
    z, x = M.water_chain(3)
    keep = np.ones(len(z), dtype=bool)
    keep[4] = False  # drop one hydrogen -> doublet
    z, x = z[keep], x[keep]
    full = O.scf(z, x, multiplicity=2, e_tol=1e-11, p_tol=1e-10, full=False)
    dc = O.dc(z, x, multiplicity=2, core_size=3, buffer_radius=500.0, kt=0.0, e_tol=1e-11, p_tol=1e-10,
              max_scf=400)
    assert dc["unrestricted"] and dc["converged"]
    assert abs(dc["total_ev"] - full["total_ev"]) < 1e-6


def test_charge_is_conserved_and_buffer_converges_monotonically():
    # tests/divide_conquer.rs:216-253 and :355-383
    z, x = M.water_chain(8)
    full = O.scf(z, x, full=False)
    errs = []
    for buf in (6.0, 9.0, 14.0):
        r = O.dc(z, x, core_size=6, buffer_radius=buf, kt=0.0)
        assert r["converged"]
        assert abs(r["charges"].sum()) < 1e-8
        errs.append(abs(r["total_ev"] - full["total_ev"]))
    assert errs[0] >= errs[1] >= errs[2]
