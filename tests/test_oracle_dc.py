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


# ------------------------------------------------------------------------------ filling (fermi.rs tests)
def _restricted(levels):
    return np.array(levels, dtype=float), np.full(len(levels), 2.0)


def test_aufbau_on_a_gapped_spectrum_is_integral():
    # fermi.rs:295-305
    e, w = _restricted([-10.0, -8.0, -6.0, 2.0, 4.0])
    f, mu, ts = O.fill(e, w, 6.0, 0.0)
    assert list(f) == [1.0, 1.0, 1.0, 0.0, 0.0]
    assert abs((w * f).sum() - 6.0) < 1e-14
    assert abs((w * f * e).sum() - 2.0 * (-10.0 - 8.0 - 6.0)) < 1e-12
    assert -6.0 < mu < 2.0
    assert ts == 0.0


def test_aufbau_shares_a_partial_shell_over_degenerate_levels():
    # fermi.rs:307-317
    e, w = _restricted([-5.0, -5.0, -5.0])
    f, _, _ = O.fill(e, w, 2.0, 0.0)
    assert np.abs(f - 1.0 / 3.0).max() < 1e-12


@pytest.mark.parametrize("kt", [0.001, 0.01, 0.1, 0.5])
def test_fermi_conserves_the_electron_count(kt):
    # fermi.rs:319-330
    e, w = _restricted([-10.0, -8.0, -7.9, -0.1, 0.0, 3.0])
    f, _, _ = O.fill(e, w, 7.0, kt)
    assert abs((w * f).sum() - 7.0) < 1e-10


def test_fermi_approaches_aufbau_as_the_temperature_falls():
    # fermi.rs:332-371
    e, w = _restricted([-10.0, -8.0, -6.0, 2.0, 4.0])
    sharp, _, _ = O.fill(e, w, 6.0, 0.0)
    prev_diff, prev_ts = np.inf, np.inf
    for kt in (0.5, 0.1, 0.01, 0.001):
        f, _, ts = O.fill(e, w, 6.0, kt)
        diff = np.abs(f - sharp).sum()
        assert diff <= prev_diff and 0.0 <= ts <= prev_ts
        prev_diff, prev_ts = diff, ts


def test_symmetric_two_level_system_has_the_closed_form_entropy():
    # two levels at -d and +d holding half their capacity: mu = 0 by symmetry and, with
    # f = 1/(1+e^{-d/kT}) below and 1-f above, T*S = -2 kT sum_w [f ln f + (1-f) ln(1-f)]
    d, kt = 0.3, 0.2
    f, mu, ts = O.fill([-d, d], [2.0, 2.0], 2.0, kt)
    lo = 1.0 / (1.0 + np.exp(-d / kt))
    assert abs(mu) < 1e-12
    assert abs(f[0] - lo) < 1e-12 and abs(f[1] - (1.0 - lo)) < 1e-12
    want = -kt * 2.0 * 2.0 * (lo * np.log(lo) + (1.0 - lo) * np.log(1.0 - lo))
    assert abs(ts - want) < 1e-12


def test_weights_carry_subsystem_factors():
    # fermi.rs:391-419: one chemical potential across differently weighted levels
    f, _, _ = O.fill([-5.0, 1.0, -4.0, -0.5], [1.0, 1.0, 1.0, 1.0], 3.0, 0.01)
    assert f[0] > 0.99 and f[2] > 0.99 and f[3] > 0.99 and f[1] < 0.01


def test_overfilling_is_an_error_and_empty_or_full_fills():
    # fermi.rs:421-438
    e, w = _restricted([-1.0, 0.0])
    with pytest.raises(ValueError):
        O.fill(e, w, 5.0, 0.0)
    with pytest.raises(ValueError):
        O.fill(e, w, -1.0, 0.0)
    f, _, _ = O.fill(e, w, 0.0, 0.1)
    assert abs((w * f).sum()) < 1e-10
    f, _, ts = O.fill(e, w, 4.0, 0.1)
    assert abs((w * f).sum() - 4.0) < 1e-10 and abs(ts) < 1e-9


def test_full_buffer_with_a_field_reproduces_the_field_scf():
    # run_divide_conquer puts the field into H_core and adds its nuclear half
    # (divide_conquer.rs:494-500); the full-buffer identity must then hold with the field on
    z, x = M.water_chain(4)
    f = [0.02, -0.01, 0.03]
    full = O.scf(z, x, e_tol=1e-11, p_tol=1e-10, full=False, electric_field=f)
    zero = O.scf(z, x, e_tol=1e-11, p_tol=1e-10, full=False)
    assert abs(full["total_ev"] - zero["total_ev"]) > 1e-3
    dc = O.dc(z, x, core_size=3, buffer_radius=500.0, kt=0.0, e_tol=1e-11, p_tol=1e-10, max_scf=400,
              electric_field=f)
    assert dc["converged"]
    assert abs(dc["total_ev"] - full["total_ev"]) < 1e-7
    assert np.abs(dc["charges"] - full["charges"]).max() < 1e-7
