"""Elements with principal quantum number n >= 4 (Zn, Ge, As, Se, Br, Sb, Te, I, Hg in AM1; Br, I in
RM1): their overlaps come from the quadrature of overlap_numeric.rs. The oracle's quadrature is
pinned against the n <= 3 closed forms (which it must reproduce, overlap.rs:717-748), the SCF of
each element's hydride must converge with a gradient equal to the derivative of its energy
(tests/element_coverage.rs:101-162); the GPU is compared with the oracle."""
import numpy as np
import pytest

from oracle import oracle as O

ANG = 1.0 / 0.529167
HEAVY_AM1 = [30, 32, 33, 34, 35, 51, 52, 53, 80]
HEAVY_RM1 = [35, 53]
# single-bond covalent radii (Cordero 2008), Angstrom, for the hydride bond lengths
RADIUS = {1: 0.31, 30: 1.22, 32: 1.20, 33: 1.19, 34: 1.20, 35: 1.20, 51: 1.39, 52: 1.38, 53: 1.39, 80: 1.32}


def hydride(z, direction=(0.41, -0.63, 0.66)):
    d = np.array(direction, dtype=float)
    d /= np.linalg.norm(d)
    bond = (RADIUS[z] + RADIUS[1]) * ANG
    return [z, 1], np.array([[0.0, 0.0, 0.0], d * bond])


def multiplicity(z, method="am1"):
    ne = O.element(z, method)["core"] + 1.0
    return 1 if round(ne) % 2 == 0 else 2


@pytest.mark.parametrize("method", ["am1", "rm1"])
def test_quadrature_reproduces_the_closed_forms(method):
    light = [z for z in range(1, 19) if O.element(z, method) is not None]
    for za in light:
        for zb in light:
            if O.element(za, method)["n"] < O.element(zb, method)["n"]:
                continue
            for r in (1.2, 2.4, 4.5):
                a = O.overlap_locals(za, zb, r, method=method)
                n = O.overlap_locals(za, zb, r, method=method, numeric=True)
                assert np.abs(a - n).max() < 1e-12, (za, zb, r)


@pytest.mark.parametrize("z,method", [(z, "am1") for z in HEAVY_AM1] + [(z, "rm1") for z in HEAVY_RM1])
def test_heavy_hydride_converges_and_gradient_is_the_energy_derivative(z, method):
    zz, x = hydride(z)
    mult = multiplicity(z, method)
    tight = dict(e_tol=1e-11, p_tol=1e-10, max_scf=500, multiplicity=mult, method=method)
    r = O.gradient(zz, x, **tight)
    assert r["status"] == 0 and r["converged"], r["message"]
    h = 1e-4
    for a in range(2):
        for k in range(3):
            xp = x.copy()
            xm = x.copy()
            xp[a, k] += h
            xm[a, k] -= h
            fd = (O.scf(zz, xp, full=False, **tight)["total_ev"] - O.scf(zz, xm, full=False, **tight)["total_ev"]) / (2 * h)
            assert abs(fd - r["gradient"][a, k]) < 5e-6


@pytest.mark.gpu
@pytest.mark.parametrize("z,method", [(z, "am1") for z in HEAVY_AM1] + [(z, "rm1") for z in HEAVY_RM1])
def test_gpu_heavy_hydrides_match_the_oracle(engine, z, method):
    zz, x = hydride(z)
    kw = dict(multiplicity=multiplicity(z, method), method=method, max_scf=500)
    o = O.gradient(zz, x, **kw)
    r = engine.gradient(zz, x, **kw)
    assert r["status"] == 0 and o["status"] == 0
    assert abs(r["total_ev"] - o["total_ev"]) <= 1e-6
    assert np.abs(r["charges"] - o["charges"]).max() <= 1e-6
    assert np.abs(r["gradient"] - o["gradient"]).max() <= 1e-6


@pytest.mark.gpu
def test_gpu_bromo_and_iodo_molecules(engine):
    """CH3Br, CH2BrI and ZnH2-like mixes through the batched kernel against the oracle."""
    mols = {
        "CH3Br": ([6, 35, 1, 1, 1], [[0, 0, 0], [0, 0, 1.94], [1.03, 0, -0.36], [-0.51, 0.89, -0.36],
                                     [-0.51, -0.89, -0.36]]),
        "CH2BrI": ([6, 35, 53, 1, 1], [[0, 0, 0], [0, 0, 1.94], [2.0, 0, -0.7], [-0.51, 0.89, -0.36],
                                       [-0.51, -0.89, -0.36]]),
        "ZnH2": ([30, 1, 1], [[0, 0, 0], [1.53, 0, 0], [-1.53, 0, 0]]),
    }
    batch = [(z, np.array(p, dtype=float) * ANG) for z, p in mols.values()]
    rb = engine.scf_batch(batch)
    for (z, x), r in zip(batch, rb):
        o = O.scf(z, x, full=False)
        assert r["status"] == 0 and o["status"] == 0
        assert abs(r["total_ev"] - o["total_ev"]) <= 1e-6
        assert np.abs(r["charges"] - o["charges"]).max() <= 1e-6
        assert abs(r["homo_ev"] - o["homo_ev"]) <= 1e-6 and abs(r["lumo_ev"] - o["lumo_ev"]) <= 1e-6
