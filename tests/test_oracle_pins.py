"""Pins on the FP64 CPU oracle (CPU only).

The oracle is the parity reference for the GPU engine, so it is checked first against what the
reference's own tests hold (MOPAC outputs, recorded heats of formation, iteration counts, known
answers) and against what the mathematics fixes (closed forms, invariants, brute force).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import molecules as M


# ------------------------------------------------------------------------------ parameters
def test_carbon_parameters_match_the_reference_table():
    # params.rs:343-352
    c = O.element(6)
    assert abs(c["uss"] - (-52.028658)) < 1e-6
    assert abs(c["gss"] - 12.23) < 1e-9
    assert c["ng"] == 4
    assert c["dd"] > 0 and c["qq"] > 0 and c["rho1"] > 0 and c["rho2"] > 0
    h = O.element(1)
    assert h["norb"] == 1 and h["dd"] == 0.0


def test_rm1_element_set_and_gaussian_counts():
    # tests/rm1.rs:57-86
    want = {1: 3, 6: 4, 7: 3, 8: 2, 9: 2, 15: 3, 16: 3, 17: 2, 35: 2, 53: 2}
    for z in range(1, 87):
        e = O.element(z, "rm1")
        if z in want:
            assert e is not None and e["ng"] == want[z], z
        else:
            assert e is None, z


@pytest.mark.parametrize("z", [5, 6, 7, 8, 9, 13, 14, 15, 16, 17])
def test_additive_terms_solve_their_defining_equations(z):
    # params.rs:283-330: rho1 = 1/(2d) with H_sp/27.21 = d/2 - 1/(2 sqrt(4 D1^2 + 1/d^2)),
    # rho2 = 1/(2q) with H_pp/27.21 = q/4 - 1/(2 sqrt(4 D2^2 + 1/q^2)) + 1/(4 sqrt(8 D2^2 + 1/q^2))
    e = O.element(z)
    d = 0.5 / e["rho1"]
    lhs = 0.5 * d - 0.5 / math.sqrt(4 * e["dd"] ** 2 + 1 / d ** 2)
    assert abs(lhs - e["hsp"] / 27.21) < 1e-12
    q = 0.5 / e["rho2"]
    hpp = max(0.5 * (e["gpp"] - e["gp2"]), 0.1) / 27.21
    lhs = 0.25 * q - 0.5 / math.sqrt(4 * e["qq"] ** 2 + 1 / q ** 2) + 0.25 / math.sqrt(8 * e["qq"] ** 2 + 1 / q ** 2)
    assert abs(lhs - hpp) < 1e-12
    assert abs(e["rho0"] - 0.5 * 27.21 / e["gss"]) < 1e-15


# ------------------------------------------------------------------------------ overlap
def _b_series(k, x):
    # B_k(x) = 2 sum_{n: k+n even} (-x)^n / (n! (k+n+1)), overlap.rs:629-645
    total, term = 0.0, 1.0
    for n in range(400):
        if (k + n) % 2 == 0:
            total += 2.0 * term / (k + n + 1)
        term *= -x / (n + 1)
        if n > 4 and abs(term) < 1e-18 * abs(total):
            break
    return total


def test_b_integrals_against_their_series():
    worst = 0.0
    for x in np.linspace(-6.0, 6.0, 301):
        if abs(x) < 1e-9:
            continue
        b = O.bintgs(x)
        for k in range(13):
            ref = _b_series(k, x)
            worst = max(worst, abs(b[k] - ref) / max(abs(ref), 1e-3))
    assert worst < 1e-12


def test_a_integrals_against_quadrature():
    from scipy.integrate import quad

    for x in [0.3, 1.7, 4.2, 9.5]:
        a = O.aintgs(x)
        for k in range(7):
            ref, _ = quad(lambda t: t ** k * math.exp(-x * t), 1.0, np.inf, epsabs=1e-14, epsrel=1e-13)
            assert abs(a[k] - ref) <= 1e-11 * abs(ref)


@pytest.mark.parametrize("r", [0.8, 1.4, 2.5, 5.0])
def test_hydrogen_ss_overlap_closed_form(r):
    # equal-exponent 1s-1s Slater overlap: S = e^{-zr} (1 + zr + (zr)^2 / 3)
    zeta = O.element(1)["zs"]
    S = O.overlap(1, [0, 0, 0], 1, [0.3 * r, -0.4 * r, math.sqrt(1 - 0.25) * r])
    zr = zeta * r
    assert abs(S[0, 0] - math.exp(-zr) * (1 + zr + zr * zr / 3)) < 1e-13


def test_overlap_rotates_as_a_tensor():
    rng = np.random.default_rng(3)
    d = np.array([1.1, -0.7, 1.9])
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    Rt = np.eye(4)
    Rt[1:, 1:] = R
    for zi, zj in [(6, 8), (8, 6), (6, 1), (17, 6), (17, 17)]:
        S = O.overlap(zi, [0, 0, 0], zj, d)
        S2 = O.overlap(zi, [0, 0, 0], zj, R @ d)
        assert np.allclose(S2, Rt @ S @ Rt.T, atol=1e-13), (zi, zj)
    # swapping the atoms transposes the block
    S = O.overlap(6, [0, 0, 0], 8, d)
    S_swap = O.overlap(8, d, 6, [0, 0, 0])
    assert np.allclose(S, S_swap.T, atol=1e-14)


# ------------------------------------------------------------------------------ integrals
def test_ssss_is_the_klopman_ohno_closed_form():
    # integrals.rs:849-858
    c = O.element(6)
    r = 2.6
    for d in ([r, 0, 0], np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8]) * r):
        w = O.pair(6, 6, d)["w"]
        assert abs(w[0, 0] - 27.21 / math.sqrt(r * r + (2 * c["rho0"]) ** 2)) < 1e-12


def _unpack(v):
    pk = [[0, 1, 3, 6], [1, 2, 4, 7], [3, 4, 5, 8], [6, 7, 8, 9]]
    return np.array([[v[pk[a][b]] for b in range(4)] for a in range(4)])


def test_two_centre_integrals_rotate_as_tensors():
    rng = np.random.default_rng(11)
    d = np.array([0.9, 1.6, -1.2])
    A = rng.normal(size=(3, 3))
    R, _ = np.linalg.qr(A)
    if np.linalg.det(R) < 0:
        R[:, 0] *= -1
    Rt = np.eye(4)
    Rt[1:, 1:] = R
    # heavy-H: the bra block transforms as R W R^T
    w1 = O.pair(8, 1, d)["w"][:, 0]
    w2 = O.pair(8, 1, R @ d)["w"][:, 0]
    assert np.allclose(_unpack(w2), Rt @ _unpack(w1) @ Rt.T, atol=1e-12)
    # heavy-heavy: T[i,j,k,l] = W[pack(i,j), pack(k,l)] is a rank-4 tensor
    pk = [[0, 1, 3, 6], [1, 2, 4, 7], [3, 4, 5, 8], [6, 7, 8, 9]]

    def full(W):
        return np.array([[[[W[pk[i][j], pk[k][l]] for l in range(4)] for k in range(4)] for j in range(4)]
                         for i in range(4)])

    T1 = full(O.pair(6, 7, d)["w"])
    T2 = full(O.pair(6, 7, R @ d)["w"])
    expect = np.einsum("ia,jb,kc,ld,abcd->ijkl", Rt, Rt, Rt, Rt, T1)
    assert np.allclose(T2, expect, atol=1e-11)


def test_integral_permutation_symmetry():
    # (mu nu | la si) of pair (A, B) equals (la si | mu nu) of pair (B, A) with the axis reversed
    d = np.array([1.2, -0.4, 2.1])
    w_ab = O.pair(6, 8, d)["w"]
    w_ba = O.pair(8, 6, -d)["w"]
    assert np.allclose(w_ab, w_ba.T, atol=1e-12)


def test_long_range_limit_is_point_charges():
    # far apart every monopole term tends to 27.21 / r
    r = 400.0
    w = O.pair(6, 8, [0, 0, r])["w"]
    assert abs(w[0, 0] * r / 27.21 - 1) < 1e-4
    assert abs(O.pair(1, 1, [r, 0, 0])["w"][0, 0] * r / 27.21 - 1) < 1e-4


# ------------------------------------------------------------------------------ linear algebra
@pytest.mark.parametrize("lapack", [True, False])
def test_eigen_known_answers_and_numpy(lapack):
    O.set_lapack(lapack)
    try:
        w, v = O.eigen(np.array([[2.0, 1.0], [1.0, 2.0]]))  # linalg.rs:479-497
        assert np.allclose(w, [1.0, 3.0], atol=1e-10)
        a = np.array([[4.0, 1.0, 2.0], [1.0, 3.0, 0.0], [2.0, 0.0, 1.0]])  # linalg.rs:499-519
        w, v = O.eigen(a)
        assert np.all(np.diff(w) >= 0)
        assert np.allclose(v @ np.diag(w) @ v.T, a, atol=1e-9)
        rng = np.random.default_rng(5)
        x = rng.normal(size=(40, 40))
        x = x + x.T
        w, v = O.eigen(x)
        wn = np.linalg.eigvalsh(x)
        assert np.allclose(w, wn, atol=1e-11)
        assert np.allclose(v.T @ v, np.eye(40), atol=1e-11)
        assert np.allclose(v @ np.diag(w) @ v.T, x, atol=1e-10)
        with pytest.raises(ValueError):
            O.eigen(np.array([[1.0, np.nan], [np.nan, 1.0]]))
    finally:
        O.set_lapack(True)


def test_diis_saddle_point():
    # linalg.rs:531-539
    c = O.diis_solve(np.array([[1.0, 0.5], [0.5, 1.0]]))
    assert abs(c[0] - 0.5) < 1e-12 and abs(c[1] - 0.5) < 1e-12 and abs(c[2] - 0.75) < 1e-12
    assert O.diis_solve(np.ones((3, 3))) is None  # dependent history -> pivot below 1e-12


def test_adiis_simplex_brute_force():
    # the projected-gradient minimiser of d.c + 1/2 c.(S + S^T).c/2 ... on the simplex, checked
    # against a grid search over the 2-simplex
    rng = np.random.default_rng(2)
    A = rng.normal(size=(3, 3))
    S = A @ A.T
    d = rng.normal(size=3)
    c = O.adiis_simplex(d, S)
    f = lambda c: 2 * d @ c + c @ S @ c
    best = min((f(np.array([a, b, 1 - a - b])), a, b) for a in np.linspace(0, 1, 201)
               for b in np.linspace(0, 1, 201) if a + b <= 1)
    assert abs(c.sum() - 1) < 1e-12 and np.all(c >= 0)
    assert f(c) <= best[0] + 1e-6


# ------------------------------------------------------------------------------ whole SCF
@pytest.mark.parametrize("name", list(M.RECORDED_HOF))
def test_recorded_heats_of_formation(name):
    # tests/auxiliary_integral_impact.rs:22-50 (recorded to four decimals)
    text, want = M.RECORDED_HOF[name]
    z, x = M.parse(text)
    r = O.scf(z, x, full=False)
    assert r["status"] == 0
    assert abs(r["heat_of_formation_kcal"] - want) < 1e-4


@pytest.mark.parametrize("method,want", [("am1", -77.16706), ("rm1", -81.09054)])
def test_mopac_co2_single_point(method, want):
    # tests/mopac_reference.rs:135-178, tolerance :74
    z, x = M.parse(M.co2(1.16))
    r = O.scf(z, x, method=method, full=False)
    assert abs(r["heat_of_formation_kcal"] - want) < 0.06


def test_formaldehyde_iteration_count_and_accelerators_agree():
    # scf.rs:1039-1040 (11 iterations under A-DIIS -> CDIIS) and scf.rs:1344-1369
    z, x = M.parse(M.FORMALDEHYDE)
    hyb = O.scf(z, x, accelerator="adiis_cdiis", full=False)
    cd = O.scf(z, x, accelerator="cdiis", full=False)
    no = O.scf(z, x, accelerator="none", full=False)
    assert hyb["iterations"] == 11
    assert abs(hyb["total_ev"] - cd["total_ev"]) < 1e-6
    assert abs(hyb["total_ev"] - no["total_ev"]) < 1e-6
    assert hyb["iterations"] < no["iterations"]


def test_water_properties_and_uhf_singlet():
    # scf.rs:1297-1312 and :1391-1418
    z, x = M.parse(M.WATER)
    r = O.scf(z, x)
    assert abs(r["heat_of_formation_kcal"] - (-59.24)) < 0.5
    assert abs(np.linalg.norm(r["dipole_debye"]) - 1.86) < 0.15
    assert abs(r["charges"].sum()) < 1e-6 and r["charges"][0] < 0 < r["charges"][1]
    u = O.scf(z, x, reference="uhf")
    assert u["unrestricted"] and abs(u["total_ev"] - r["total_ev"]) < 1e-6


def test_methyl_radical_uhf():
    # scf.rs:1372-1388
    z, x = M.parse(M.METHYL)
    r = O.scf(z, x, multiplicity=2)
    assert r["status"] == 0 and r["unrestricted"]
    assert abs(r["heat_of_formation_kcal"] - 32.0) < 6.0
    bad = O.scf(z, x, multiplicity=2, reference="rhf", full=False)
    assert bad["status"] == 1


def test_scf_invariants():
    # tests/theory_components.rs:458-601: P^2 = 2P, Tr P = N, [F,P] -> 0, E = 1/2 Tr P(H+F)
    z, x = M.parse(M.FORMALDEHYDE)
    r = O.scf(z, x)
    P = r["density"]
    assert np.allclose(P @ P, 2 * P, atol=1e-9)
    assert abs(np.trace(P) - 12.0) < 1e-9
    H, _ = O.core(z, x)
    F = O.fock(z, x, P)
    assert np.allclose(F, F.T, atol=1e-12)
    assert np.linalg.norm(F @ P - P @ F) < 1e-6
    e = 0.5 * np.sum(P * (H + F))
    assert abs(e - r["electronic_ev"]) < 1e-11 * abs(e)


def test_rigid_motion_invariance():
    # tests/theory_components.rs:634-691
    z, x = M.parse(M.FORMALDEHYDE)
    a = O.scf(z, x, full=False)
    th = 0.7
    R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    b = O.scf(z, x @ R.T + np.array([3.0, -1.0, 2.0]), full=False)
    assert abs(a["total_ev"] - b["total_ev"]) < 1e-8
    assert np.allclose(a["charges"], b["charges"], atol=1e-8)


def test_error_paths():
    z, x = M.parse(M.WATER)
    assert O.scf(z, x, max_scf=3, full=False)["status"] == 4  # scf_hardening.rs:114-136
    assert O.scf(z, x, multiplicity=2, full=False)["status"] == 1  # parity
    assert O.scf(np.array([2, 1, 1]), x, full=False)["status"] == 2  # He is not parameterized
    bad = x.copy()
    bad[1, 0] = np.nan
    assert O.scf(z, bad, full=False)["status"] == 1


def test_planted_mistake_is_caught():
    """A wrong sign in one exchange term must move the energy far past the parity tolerance:
    the tolerances used by the GPU tests are tight enough to see an error of this kind."""
    z, x = M.parse(M.WATER)
    r = O.scf(z, x)
    H, _ = O.core(z, x)
    P = r["density"]
    F = O.fock(z, x, P)
    F_wrong = O.fock(z, x, P, spin_scale=-0.5)
    e_ok = 0.5 * np.sum(P * (H + F))
    e_wrong = 0.5 * np.sum(P * (H + F_wrong))
    assert abs(e_ok - e_wrong) > 1.0


# ------------------------------------------------------------------------------ external field
def test_zero_field_changes_nothing():
    # tests/external_field.rs:48-55
    z, x = M.parse(M.WATER)
    bare = O.scf(z, x, full=False)
    zero = O.scf(z, x, full=False, electric_field=[0.0, 0.0, 0.0])
    assert abs(bare["total_ev"] - zero["total_ev"]) < 1e-12
    assert zero["field_nuclear_ev"] == 0.0


@pytest.mark.parametrize("name", ["water", "formaldehyde"])
def test_field_derivative_is_minus_the_dipole(name):
    """tests/external_field.rs:61-90: dE/dF_alpha = -mu_alpha (Hellmann-Feynman; E is variational
    in P). Central differences of the oracle's energy against its own dipole, in e*Bohr; this
    fixes the sign and size of both the electronic (H_core) and the nuclear halves."""
    z, x = M.parse(M.WATER if name == "water" else M.FORMALDEHYDE)
    tight = dict(full=False, e_tol=1e-12, p_tol=1e-11)
    r0 = O.scf(z, x, **tight)
    mu = r0["dipole_debye"] / 2.541746473  # e*Bohr
    h = 1e-4  # eV per (e*Bohr)
    for a in range(3):
        f = np.zeros(3)
        f[a] = h
        ep = O.scf(z, x, electric_field=f, **tight)["total_ev"]
        em = O.scf(z, x, electric_field=-f, **tight)["total_ev"]
        assert (ep - em) / (2 * h) == pytest.approx(-mu[a], abs=2e-6)


def test_field_nuclear_term():
    # -F . sum_a Z_a R_a (dipole.rs:146-153), with core charges 6 (O) and 1 (H)
    z, x = M.parse(M.WATER)
    f = np.array([0.3, -0.2, 0.1])
    r = O.scf(z, x, full=False, electric_field=f)
    zc = np.array([6.0 if v == 8 else 1.0 for v in z])
    assert r["field_nuclear_ev"] == pytest.approx(-f @ (zc[:, None] * x).sum(axis=0), rel=1e-14)
    assert r["total_ev"] == pytest.approx(r["electronic_ev"] + r["core_ev"] + r["field_nuclear_ev"], abs=1e-9)
