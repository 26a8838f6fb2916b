"""The device symmetric eigensolver (``am1g_symmetric_eigen``: Householder tridiagonalization,
tridiagonal divide and conquer, blocked back-transformation) against numpy.linalg.eigh, on random
matrices, exactly degenerate spectra (heavy deflation) and the reference's known answers
(linalg.rs:479-519). Tolerances scale with the matrix norm and the order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check(engine, a, tol=1e-12):
    w, v = engine.eigh(a)
    n = a.shape[0]
    scale = max(1.0, np.abs(a).max() * n ** 0.5)
    wr = np.linalg.eigvalsh(a)
    assert np.all(np.diff(w) >= 0.0)
    assert np.abs(w - wr).max() <= tol * scale
    assert np.abs(a @ v - v * w).max() <= 10 * tol * scale
    assert np.abs(v.T @ v - np.eye(n)).max() <= 10 * tol * max(1.0, n / 100)
    return w, v


def test_known_answers(engine):
    # linalg.rs:479-519
    w, v = engine.eigh(np.array([[2.0, 1.0], [1.0, 2.0]]))
    assert np.allclose(w, [1.0, 3.0], atol=1e-12)
    a = np.array([[4.0, 1.0, 0.5], [1.0, 3.0, 0.2], [0.5, 0.2, 1.0]])
    w, v = engine.eigh(a)
    assert np.abs(v @ np.diag(w) @ v.T - a).max() < 1e-12
    w, _ = engine.eigh(np.array([[5.0]]))
    assert w[0] == 5.0


@pytest.mark.parametrize("n", [3, 17, 64, 65, 100, 129, 257, 600])
def test_random_symmetric(engine, n):
    rng = np.random.default_rng(n)
    b = rng.normal(size=(n, n))
    _check(engine, 0.5 * (b + b.T))


def test_order_above_the_register_resident_reflector_step(engine):
    # the first columns of a 4113-order matrix (more than 4096 trailing rows) take the
    # 256-thread next-reflector kernel, the rest the register-resident 1024-thread one
    rng = np.random.default_rng(4113)
    b = rng.normal(size=(4113, 4113))
    _check(engine, 0.5 * (b + b.T))


def test_exactly_degenerate_spectrum(engine):
    # block diagonal with repeated blocks, rotated: every eigenvalue 4-fold degenerate
    rng = np.random.default_rng(3)
    blk = rng.normal(size=(50, 50))
    blk = blk + blk.T
    a = np.kron(np.eye(4), blk)
    q, _ = np.linalg.qr(rng.normal(size=(200, 200)))
    _check(engine, q @ a @ q.T)


def test_diagonal_and_clustered(engine):
    rng = np.random.default_rng(5)
    d = np.sort(rng.normal(size=300))
    _check(engine, np.diag(d))                      # fully deflated merges
    bands = np.concatenate([-30 + 1e-3 * rng.random(150), -10 + 1e-6 * rng.random(150), 5 + rng.random(100)])
    q, _ = np.linalg.qr(rng.normal(size=(400, 400)))
    _check(engine, (q * bands) @ q.T)               # tight clusters, as in large water clusters


def test_water_cluster_fock_size(engine):
    """An order-3072 matrix with a Fock-like spectrum: the 512-water workload's eigensolve size."""
    rng = np.random.default_rng(7)
    n = 3072
    b = rng.normal(size=(n, n)) / np.sqrt(n)
    a = 0.5 * (b + b.T) + np.diag(np.linspace(-40.0, 10.0, n))
    _check(engine, a)


def test_non_finite_input_is_a_linear_algebra_error(engine):
    a = np.eye(4)
    a[1, 2] = a[2, 1] = np.nan
    with pytest.raises(ValueError, match="LinearAlgebra"):
        engine.eigh(a)


@pytest.mark.parametrize("n", [8, 63, 257, 600, 1100])
def test_blocked_tridiagonalization(n):
    """The panel-blocked tridiagonalization (tunable ``tri_blocked_min``) on its own engine:
    random matrices, a spectrum with exact 4-fold degeneracy (heavy deflation) and a diagonal
    matrix, against numpy, and against the column-at-a-time reduction of the same matrix."""
    import am1g

    eng = am1g.Engine(0)
    try:
        eng.set_tunable("tri_blocked_min", 8)
        rng = np.random.default_rng(1000 + n)
        b = rng.normal(size=(n, n))
        a = 0.5 * (b + b.T)
        w, _ = _check(eng, a)
        q, _ = np.linalg.qr(rng.normal(size=(n, n)))
        lam = np.repeat(np.arange(n // 4 + 1, dtype=float), 4)[:n]
        _check(eng, (q * lam) @ q.T)
        _check(eng, np.diag(rng.normal(size=n)))
        eng.set_tunable("tri_blocked_min", 1 << 30)
        w2, _ = eng.eigh(a)
        assert np.abs(w - w2).max() <= 1e-11 * max(1.0, np.abs(w).max())
    finally:
        eng.close()
