"""Far-field monopole treatment (Am1Options::multipole_cutoff, farfield.rs:80-236): pairs beyond
the cutoff leave the pair list (no integrals, no H_core terms, no core-core term) and interact
through their net charges, F_aa -= V_a with V_a = sum_far gamma_ab Q_b, and the final energy
gains 1/2 sum_a Z_a V_a (fock.rs:145-232, 278-318; scf.rs:699-701).

The oracle is pinned by the reference's own checks (tests/farfield.rs): a zero or absent cutoff
changes nothing bit for bit (:149-184), the error falls as the cutoff grows and vanishes once no
pair is far (:64-106), and the screened force is the derivative of the screened energy
(:109-147), which a wrong sign or a missing factor in either the Fock shift or the energy term
breaks (the SCF energy would no longer be stationary). The GPU paths (the batched kernel, the
large-system path, D&C and the gradient) are compared with the oracle."""
import numpy as np
import pytest

from oracle import oracle as O

ANG = 1.0 / 0.529167
KW = dict(e_tol=1e-9, p_tol=1e-8, max_scf=400, full=False)


def water_cluster(n_waters, spacing=4.0):
    """`n` water molecules on a cubic grid (tests/farfield.rs:20-50)."""
    side = 1
    while side ** 3 < n_waters:
        side += 1
    z, x = [], []
    made = 0
    for i in range(side):
        for j in range(side):
            for k in range(side):
                if made >= n_waters:
                    break
                shift = np.array([i, j, k], dtype=float) * spacing * ANG
                for zz, r in ((8, (0.0, 0.0, 0.0)), (1, (0.9614, 0.0, 0.0)), (1, (-0.2246, 0.9348, 0.0))):
                    z.append(zz)
                    x.append(np.array(r) * ANG + shift)
                made += 1
    return np.array(z), np.array(x)


def test_zero_or_absent_cutoff_changes_nothing():
    z, x = water_cluster(8)
    a = O.scf(z, x, **KW)
    for cut in (None, 0.0, -3.0, float("nan"), 1.0e6):
        b = O.scf(z, x, multipole_cutoff=cut, **KW)
        assert b["total_ev"] == a["total_ev"], cut
        assert b["iterations"] == a["iterations"]


def test_error_falls_as_the_cutoff_grows():
    z, x = water_cluster(27)
    exact = O.scf(z, x, **KW)
    assert exact["status"] == 0
    errs = []
    for cut in (10.0, 15.0, 20.0, 45.0):
        r = O.scf(z, x, multipole_cutoff=cut, **KW)
        assert r["status"] == 0, cut
        errs.append(abs(r["total_ev"] - exact["total_ev"]) / len(z))
    assert errs[0] > errs[1] > errs[2], errs
    assert errs[0] > 1e-4  # the approximation is really on at 10 Bohr
    assert errs[-1] < 1e-4  # no pair beyond 45 Bohr here: exact


def test_screened_force_is_the_derivative_of_the_screened_energy():
    z, x = water_cluster(8)
    cut = 12.0
    g = O.gradient(z, x, multipole_cutoff=cut, e_tol=1e-12, p_tol=1e-12, max_scf=400)["gradient"]
    h = 1e-4
    worst = 0.0
    for a in (0, 4, 9, 17, 23):
        for k in range(3):
            xp = x.copy()
            xm = x.copy()
            xp[a, k] += h
            xm[a, k] -= h
            ep = O.scf(z, xp, multipole_cutoff=cut, full=False, e_tol=1e-12, p_tol=1e-12, max_scf=400)["total_ev"]
            em = O.scf(z, xm, multipole_cutoff=cut, full=False, e_tol=1e-12, p_tol=1e-12, max_scf=400)["total_ev"]
            worst = max(worst, abs((ep - em) / (2 * h) - g[a, k]))
    assert worst < 1e-5, worst


def test_far_pairs_add_the_monopole_energy():
    """Two neutral-ish fragments far apart: with every inter-fragment pair beyond the cutoff,
    the energy equals the two fragments' SCF at fixed geometry plus (to the converged charges)
    sum over inter-fragment pairs of Q_a Q_b gamma_ab, the monopole energy itself."""
    z1, x1 = water_cluster(1)
    x2 = x1 + np.array([0.3, 14.0, -0.4]) * ANG
    z = np.concatenate([z1, z1])
    x = np.concatenate([x1, x2])
    cut = 10.0 * ANG
    r = O.scf(z, x, multipole_cutoff=cut, full=True, **{k: v for k, v in KW.items() if k != "full"})
    q = r["charges"]
    gam = 0.0
    for a in range(3):
        for b in range(3, 6):
            rho = O.element(int(z[a]))["rho0"] + O.element(int(z[b]))["rho0"]
            d = np.linalg.norm(x[b] - x[a])
            gam += 27.21 / np.sqrt(d * d + rho * rho) * q[a] * q[b]
    single = O.scf(z1, x1, **KW)
    # each fragment sits in the other's monopole field, which polarizes it slightly; the
    # energy is stationary, so the change is second order in that field (3e-7 eV here, against a
    # monopole energy of -2.6e-4 eV)
    assert abs(gam) > 1e-4
    assert abs(r["total_ev"] - (2 * single["total_ev"] + gam)) < 1e-5


def test_dc_full_buffer_reproduces_the_screened_scf():
    """The D&C full-buffer identity (tests/divide_conquer.rs:160-214) with the far field on in
    both: the D&C builds its Fock matrices from the same screened core (divide_conquer.rs:428)."""
    from tests import molecules as M

    z, x = M.water_chain(6)
    cut = 9.0
    full = O.scf(z, x, multipole_cutoff=cut, e_tol=1e-11, p_tol=1e-10, full=False)
    exact = O.scf(z, x, e_tol=1e-11, p_tol=1e-10, full=False)
    assert abs(full["total_ev"] - exact["total_ev"]) > 1e-5  # some pairs are far
    dc = O.dc(z, x, core_size=3, buffer_radius=500.0, kt=0.0, e_tol=1e-11, p_tol=1e-10, max_scf=400,
              multipole_cutoff=cut)
    assert dc["converged"]
    assert abs(dc["total_ev"] - full["total_ev"]) < 1e-7
    assert np.abs(dc["charges"] - full["charges"]).max() < 1e-7


# ---------------------------------------------------------------------------------- GPU parity
E_TOL, Q_TOL = 1e-6, 1e-6


def _radical(n_waters):
    z, x = water_cluster(n_waters)
    keep = np.ones(len(z), dtype=bool)
    keep[2] = False  # one OH radical
    return z[keep], x[keep]


def _cmp(r, o):
    assert r["status"] == 0 and o["status"] == 0, r.get("message")
    assert abs(r["total_ev"] - o["total_ev"]) <= E_TOL, (r["total_ev"], o["total_ev"])
    assert np.abs(np.asarray(r["charges"]) - np.asarray(o["charges"])).max() <= Q_TOL
    assert abs(r["iterations"] - o["iterations"]) <= 1


@pytest.mark.gpu
def test_gpu_batched_kernel(engine):
    cases = [(water_cluster(20), 1, 10.0), (water_cluster(20), 1, 15.0), (_radical(12), 2, 9.0),
             (water_cluster(8), 1, None)]
    kws = [dict(multiplicity=m, multipole_cutoff=c) for _, m, c in cases]
    mols = [(z, x, kw) for ((z, x), _, _), kw in zip(cases, kws)]
    res = engine.scf_batch(mols)
    for (z, x, kw), r in zip(mols, res):
        o = O.scf(z, x, full=False, **kw)
        _cmp(r, o)


@pytest.mark.gpu
def test_gpu_large_path(engine, monkeypatch):
    monkeypatch.setenv("AM1G_FORCE_LARGE_PATH", "1")
    for (z, x), mult, cut in ((water_cluster(27), 1, 10.0), (_radical(12), 2, 9.0)):
        r = engine.scf(z, x, full=False, multiplicity=mult, multipole_cutoff=cut)
        o = O.scf(z, x, full=False, multiplicity=mult, multipole_cutoff=cut)
        _cmp(r, o)


@pytest.mark.gpu
def test_gpu_divide_and_conquer(engine):
    z, x = water_cluster(27)
    kw = dict(core_size=3, buffer_radius=9.0, kt=0.05, e_tol=1e-8, p_tol=1e-7, max_scf=300, multipole_cutoff=12.0)
    r = engine.dc(z, x, **kw)
    o = O.dc(z, x, **kw)
    assert r["status"] == 0 and o["status"] == 0
    assert abs(r["total_ev"] - o["total_ev"]) <= E_TOL
    assert np.abs(np.asarray(r["charges"]) - o["charges"]).max() <= Q_TOL


@pytest.mark.gpu
def test_gpu_gradient(engine):
    z, x = water_cluster(8)
    for kw in (dict(multipole_cutoff=12.0), dict(multipole_cutoff=9.0, multiplicity=1)):
        r = engine.gradient(z, x, e_tol=1e-10, p_tol=1e-9, **kw)
        o = O.gradient(z, x, e_tol=1e-10, p_tol=1e-9, **kw)
        assert r["status"] == 0
        assert np.abs(r["gradient"] - o["gradient"]).max() < 1e-5
