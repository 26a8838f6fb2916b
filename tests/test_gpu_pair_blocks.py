"""Per-block parity of the pair-integral kernel (k_pairs) and of the core Hamiltonian the device
builds from it (am1g_core_hamiltonian, through the C ABI), against the oracle's
pair_two_electron (integrals.rs:207-277) and build_core_with_neighbors / core-core energy
(hamiltonian.rs:211-366, repulsion.rs:38-188).

Every element pair of both parameter sets meets here at several separations (0.8 to 12 A, both
sides of the overlap's |x| <= 5 series switch, heavy-heavy, heavy-hydrogen and hydrogen-hydrogen
classes, arbitrary orientations), so a wrong term in a rarely used pair (F, P, Cl, the n >= 4
elements, RM1) fails a block here even where it would hide under the SCF's 1e-6 eV energy test.
Tolerance: 1e-12 relative to the block's largest entry (the two codes evaluate the same
closed forms in a different order)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ANG = 1.0 / 0.529167
REL = 1e-12
METHOD_ELEMENTS = {
    "am1": [z for z in range(1, 90) if O.element(z, "am1") is not None],
    "rm1": [z for z in range(1, 90) if O.element(z, "rm1") is not None],
}


def _multiplicity(z, method):
    ne = sum(O.element(int(a), method)["core"] for a in z)
    return 1 if round(ne) % 2 == 0 else 2


def _scatter(zs, rng, spacing):
    """Atoms on a jittered cubic grid (no two closer than ~0.6 spacing), shuffled order."""
    n = len(zs)
    side = int(np.ceil(n ** (1 / 3)))
    cells = np.array([(i, j, k) for i in range(side) for j in range(side) for k in range(side)], dtype=float)
    cells = cells[rng.permutation(len(cells))[:n]]
    x = (cells + rng.uniform(-0.2, 0.2, cells.shape)) * spacing * ANG
    return np.array(zs, dtype=np.int64), x


def _check(engine, z, x, method):
    mult = _multiplicity(z, method)
    dev = engine.core_hamiltonian(z, x, method=method, multiplicity=mult)
    worst = 0.0
    for (a, b), w in dev["pairs"].items():
        ref = O.pair(int(z[a]), int(z[b]), x[b] - x[a], method=method)["w"]
        assert w.shape == ref.shape, (z[a], z[b])
        scale = max(np.abs(ref).max(), 1e-300)
        err = np.abs(w - ref).max() / scale
        worst = max(worst, err)
        assert err <= REL, (method, int(z[a]), int(z[b]), float(np.linalg.norm(x[b] - x[a]) / ANG), err)
    H, ecore = O.core(z, x, method=method)
    assert dev["h_core"].shape == H.shape
    herr = np.abs(dev["h_core"] - H).max() / np.abs(H).max()
    assert herr <= REL, herr
    assert abs(dev["core_ev"] - ecore) <= REL * max(1.0, abs(ecore))
    return worst, herr


@pytest.mark.parametrize("method", ["am1", "rm1"])
@pytest.mark.parametrize("spacing", [1.6, 3.0, 6.0])
def test_every_element_pair_block(engine, method, spacing):
    rng = np.random.default_rng(int(spacing * 10) + (7 if method == "rm1" else 0))
    zs = METHOD_ELEMENTS[method] + [1, 1, 1]  # three hydrogens: H-H pairs and H against every element
    z, x = _scatter(zs, rng, spacing)
    worst, herr = _check(engine, z, x, method)
    print(f"{method} spacing {spacing} A: {len(z)} atoms, worst block error {worst:.2e}, H_core {herr:.2e}")


@pytest.mark.parametrize("method", ["am1", "rm1"])
def test_distance_scan_of_each_pair_class(engine, method):
    """Two atoms along a random axis from 0.8 to 12 A: the overlap's series and recurrence
    branches and the integrals' short- and long-range regimes."""
    rng = np.random.default_rng(3)
    els = METHOD_ELEMENTS[method]
    heavy = [e for e in els if e != 1]
    combos = [(1, 1)] + [(h, 1) for h in heavy] + [(a, b) for a in heavy for b in heavy if a <= b][::3]
    for za, zb in combos:
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        for r in (0.8, 1.3, 2.1, 3.4, 5.5, 12.0):
            x = np.array([[0.1, -0.2, 0.3], [0.1, -0.2, 0.3]]) * ANG
            x[1] += d * r * ANG
            _check(engine, np.array([za, zb]), x, method)


def test_field_enters_h_core(engine):
    """The external field's dipole operator is part of the exported H_core: its change equals
    the oracle's field-free H_core difference pattern (diagonal F.R_a, s-p entries F.dd)."""
    z = np.array([8, 1, 1])
    x = np.array([[0.0, 0.0, 0.0], [0.9614, 0.0, 0.0], [-0.2246, 0.9348, 0.0]]) * ANG
    f = np.array([0.01, -0.02, 0.015])
    h0 = engine.core_hamiltonian(z, x)["h_core"]
    h1 = engine.core_hamiltonian(z, x, electric_field=f)["h_core"]
    dd = O.element(8, "am1")["dd"]
    expect = np.zeros_like(h0)
    aoff = [0, 4, 5]
    for a, off in enumerate(aoff):
        n = 4 if z[a] != 1 else 1
        for k in range(n):
            expect[off + k, off + k] = f @ x[a]
    for k in range(3):
        expect[0, 1 + k] = expect[1 + k, 0] = f[k] * dd
    assert np.abs((h1 - h0) - expect).max() < 1e-12


def test_short_buffer_is_refused(engine):
    import ctypes

    import am1g

    z = np.array([6, 8, 1, 1])
    x = np.array([[0, 0, 0], [1.2, 0, 0], [-0.6, 0.9, 0], [-0.6, -0.9, 0]], dtype=float) * ANG
    keep = []
    rq = engine._request(z, x, keep)
    w = np.zeros(5)
    st = engine._L.am1g_core_hamiltonian(engine._ctx, ctypes.byref(rq), None, None,
                                         w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 5)
    assert am1g.STATUS_NAMES[st] == "InvalidInput"
    assert "pair_blocks" in engine.last_error()
