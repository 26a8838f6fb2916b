"""The device GEMM (``am1g_matmul``: FP64 matrix cores, or FP32 with FP32 accumulation) against
numpy, in the forms of the reference's ``Matrix`` products (linalg.rs:92-181): plain,
transposed operands and the accumulating ``dst += alpha (...)`` forms. The shapes of the
reference's own check of the transpose-free products (linalg.rs:391-422) come first, then
shapes that are not multiples of the 64 x 64 tile, an empty inner dimension, and the skinny
products of the eigensolver's back-transformation, which split the inner dimension over
several workgroups."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def filled(rows, cols, seed):
    """The reference test-suite's deterministic fill (linalg.rs:367-377): a 64-bit LCG, values in
    [-0.5, 0.5)."""
    mask = (1 << 64) - 1
    s = (seed * 6364136223846793005 + 1) & mask
    out = np.empty(rows * cols)
    for i in range(rows * cols):
        s = (s * 6364136223846793005 + 1442695040888963407) & mask
        out[i] = (s >> 11) / float(1 << 53) - 0.5
    return out.reshape(rows, cols)


@pytest.mark.parametrize("m,k,n", [(7, 5, 9), (64, 32, 48), (33, 65, 17)])
def test_transpose_free_products(engine, m, k, n):
    # linalg.rs:391-422: the transposed-operand products equal the products of the materialized
    # transposes; the accumulating form cancels its own product
    a = filled(m, k, 3)
    b = filled(n, k, 5)
    c = filled(k, n, 11)
    d = filled(k, m, 13)
    want = a @ b.T
    assert np.abs(engine.matmul(a, b, trans_b=True) - want).max() < 1e-14
    assert np.abs(engine.matmul(a, b.T.copy()) - want).max() < 1e-14
    want2 = d.T @ c
    assert np.abs(engine.matmul(d, c, trans_a=True) - want2).max() < 1e-14
    acc = engine.matmul(a, b, c=want, trans_b=True, alpha=-1.0, beta=1.0)
    assert np.abs(acc).max() < 1e-14


@pytest.mark.parametrize("ta", [False, True])
@pytest.mark.parametrize("tb", [False, True])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (130, 70, 200), (257, 129, 63), (200, 300, 1000), (64, 3072, 2900)])
def test_shapes_fp64(engine, ta, tb, m, n, k):
    rng = np.random.default_rng(m * 7919 + n * 31 + k)
    a = rng.normal(size=(k, m) if ta else (m, k))
    b = rng.normal(size=(n, k) if tb else (k, n))
    c0 = rng.normal(size=(m, n))
    ref = 0.75 * ((a.T if ta else a) @ (b.T if tb else b)) - 0.5 * c0
    got = engine.matmul(a, b, c=c0, trans_a=ta, trans_b=tb, alpha=0.75, beta=-0.5)
    # FP64 with k-term sums: rounding grows like sqrt(k) eps |a||b|
    tol = 64 * np.finfo(float).eps * np.sqrt(k) * 4.0
    assert np.abs(got - ref).max() <= tol * max(1.0, np.abs(ref).max())


def test_empty_inner_dimension(engine):
    c0 = np.arange(12.0).reshape(3, 4)
    got = engine.matmul(np.zeros((3, 0)), np.zeros((0, 4)), c=c0, beta=2.0)
    assert np.array_equal(got, 2.0 * c0)
    got = engine.matmul(np.zeros((3, 0)), np.zeros((0, 4)), c=c0)
    assert np.array_equal(got, np.zeros((3, 4)))


def test_beta_zero_ignores_nan(engine):
    # BLAS semantics: beta = 0 does not read c, so NaN in c does not reach the result
    a = filled(40, 30, 1)
    b = filled(30, 50, 2)
    got = engine.matmul(a, b, c=np.full((40, 50), np.nan))
    assert np.abs(got - a @ b).max() < 1e-13


def test_density_product_fp32(engine):
    # the mixed-precision density 2 C_occ C_occ^T: FP32 operands and accumulation
    rng = np.random.default_rng(5)
    c = np.linalg.qr(rng.normal(size=(600, 600)))[0][:, :400]
    got = engine.matmul(c, c, trans_b=True, alpha=2.0, precision="mixed")
    cf = c.astype(np.float32).astype(np.float64)
    ref = 2.0 * cf @ cf.T
    # FP32 accumulation over 400 terms of size <= 1/400 each
    assert np.abs(got - ref).max() < 2e-5
    assert np.abs(got - 2.0 * c @ c.T).max() < 2e-5


def test_determinism(engine):
    # split-K partials are summed in a fixed order: the same inputs give the same bits
    rng = np.random.default_rng(9)
    a = rng.normal(size=(3000, 64))
    b = rng.normal(size=(3000, 2000))
    r1 = engine.matmul(a, b, trans_a=True)
    r2 = engine.matmul(a, b, trans_a=True)
    assert np.array_equal(r1, r2)
    assert np.abs(r1 - a.T @ b).max() < 1e-11
