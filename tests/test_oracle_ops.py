"""Pins of the oracle's sparse building blocks against the mathematics:
the conservative product (src/common/state_manager.cpp:60; residual form
src/common/dropping.cpp:97-104) against exact rational products and Boolean
pattern products, and the 2x2 least-squares solve
(src/common/sparse_operations.cpp:58-73) against closed forms and numpy's
minimum-norm lstsq."""
from fractions import Fraction

import numpy as np
import pytest
from scipy import sparse


def rand_csc(m, n, density, seed, zeros=0):
    rng = np.random.default_rng(seed)
    A = sparse.random(m, n, density=density, random_state=rng, format="csc")
    A.data = rng.standard_normal(A.nnz)
    if zeros:
        A.data[rng.choice(A.nnz, min(zeros, A.nnz), replace=False)] = 0.0  # explicit stored zeros
    A.sort_indices()
    return A


def boolean_pattern(X, Y):
    bx = sparse.csc_matrix((np.ones(X.nnz), X.indices, X.indptr), shape=X.shape)
    by = sparse.csc_matrix((np.ones(Y.nnz), Y.indices, Y.indptr), shape=Y.shape)
    P = (bx @ by).tocsc()
    P.sort_indices()
    return P


def exact_product(X, Y):
    Xd = X.toarray()
    Yd = Y.toarray()
    out = {}
    for j in range(Yd.shape[1]):
        for i in range(Xd.shape[0]):
            s = Fraction(0)
            for k in range(Xd.shape[1]):
                if Xd[i, k] != 0 and Yd[k, j] != 0:
                    s += Fraction(Xd[i, k]) * Fraction(Yd[k, j])
            out[i, j] = s
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_product_pattern_and_values(oracle_lib, seed):
    X = rand_csc(13, 11, 0.3, seed, zeros=3)
    Y = rand_csc(11, 9, 0.35, seed + 10, zeros=2)
    C = oracle_lib.product(X, Y)
    P = boolean_pattern(X, Y)
    # structure: exactly the Boolean product of the stored patterns (explicit zeros kept)
    np.testing.assert_array_equal(C.indptr, P.indptr)
    np.testing.assert_array_equal(C.indices, P.indices)
    ex = exact_product(X, Y)
    for j in range(C.shape[1]):
        for p in range(C.indptr[j], C.indptr[j + 1]):
            i = C.indices[p]
            e = float(ex[i, j])
            assert abs(C.data[p] - e) <= 16 * np.finfo(float).eps * max(abs(e), 1e-300) + 1e-300 or C.data[p] == e


def test_product_is_not_transposed(oracle_lib):
    X = rand_csc(8, 8, 0.4, 5)
    Y = rand_csc(8, 8, 0.4, 6)
    C = oracle_lib.product(X, Y).toarray()
    np.testing.assert_allclose(C, X.toarray() @ Y.toarray(), rtol=1e-14, atol=1e-14)
    assert not np.allclose(C, Y.toarray() @ X.toarray())


def test_residual_form_and_zero_pruning(oracle_lib):
    # integer values: A @ M has exact zeros by cancellation
    A = sparse.csc_matrix(np.array([[2.0, -1.0, 0.0], [-1.0, 2.0, -1.0], [0.0, -1.0, 2.0]]))
    M = sparse.csc_matrix(np.array([[1.0, 1.0, 0.0], [1.0, 2.0, 0.0], [0.0, 0.0, 0.5]]))
    R = oracle_lib.product(A, M, residual=True)
    np.testing.assert_array_equal(R.toarray(), np.eye(3) - A.toarray() @ M.toarray())
    # (A M)[0,1] = 2 - 2 = 0 is stored without pruning and dropped with it
    kept = oracle_lib.product(A, M, residual=True, prune_zero=False)
    pruned = oracle_lib.product(A, M, residual=True, prune_zero=True)
    assert kept.nnz > pruned.nnz
    assert np.all(pruned.data != 0.0)
    np.testing.assert_array_equal(pruned.toarray(), kept.toarray())
    # the diagonal is present in the residual even where A M has no entry
    # (columns 0 and 1 of A Z are empty; the reference rejects nnz = 0 inputs)
    Z = sparse.csc_matrix(([0.5], ([0], [2])), shape=(3, 3))
    R0 = oracle_lib.product(A, Z, residual=True)
    np.testing.assert_array_equal(R0.toarray(), np.eye(3) - A.toarray() @ Z.toarray())
    assert R0[0, 0] == 1.0 and R0[1, 1] == 1.0 and (R0[:, 0].nnz, R0[:, 1].nnz) == (1, 1)


@pytest.mark.parametrize("seed", range(6))
def test_ls2x2_nonsingular_matches_solve(oracle_lib, seed):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((5, 2))
    G = B.T @ B
    x = rng.standard_normal(2)
    got = oracle_lib.solve_ls2x2(G, x)
    np.testing.assert_allclose(got, np.linalg.solve(G, x), rtol=1e-12, atol=1e-14)


def scaled_min_norm(G, x):
    """Minimum-norm solution in the coordinates scaled by 1/sqrt|diag(G)|
    (the reference scales before its complete orthogonal decomposition)."""
    d = np.sqrt(np.abs(np.diag(G)))
    s = np.where(d > 0, 1.0 / np.where(d > 0, d, 1.0), 1.0)
    y = np.linalg.lstsq(s[:, None] * G * s[None, :], s * x, rcond=None)[0]
    return s * y


def test_ls2x2_rank_one_is_scaled_minimum_norm(oracle_lib):
    v = np.array([3.0, -2.0])
    G = np.outer(v, v)  # rank 1
    x = np.array([1.0, -2.0 / 3.0])  # in range(G)
    got = oracle_lib.solve_ls2x2(G, x)
    np.testing.assert_allclose(got, scaled_min_norm(G, x), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(got, [1.0 / 18.0, -1.0 / 12.0], rtol=1e-14)
    # Gram of two collinear directions a and 2a
    a = np.array([1.0, 2.0, 2.0])
    G = np.array([[a @ a, 2 * a @ a], [2 * a @ a, 4 * a @ a]])
    x = np.array([a @ a, 2 * a @ a])
    got = oracle_lib.solve_ls2x2(G, x)
    np.testing.assert_allclose(got, scaled_min_norm(G, x), rtol=1e-12, atol=1e-14)
    # either way the combination solves the least-squares problem
    np.testing.assert_allclose(G @ got, x, rtol=1e-12)


@pytest.mark.parametrize("rho", [1e-5, 1e-7])
def test_ls2x2_rank_eps_takes_small_pivots_as_zero(oracle_lib, rho):
    """The FP32 inner-outer LOMR form solves its per-column 2x2 Grams with the
    rank threshold rank_eps = sqrt(FLT_EPSILON) instead of DBL_EPSILON.  Two
    directions a and q = t a + rho e (e orthogonal to a) give a Gram whose
    scaled second pivot is about (rho |e| / (t |a|))^2, between the two
    thresholds here.  With sqrt(FLT_EPSILON) the solve is the scaled
    minimum-norm solution of the exact rank-1 Gram (rho = 0) up to O(rho);
    with DBL_EPSILON it is the full-rank solution, which resolves the e
    component of r with the coefficient gamma = (r.e) / (rho e.e): large, and
    cancelled in delta."""
    a = np.array([1.0, 2.0, 2.0, -1.0])
    e = np.array([2.0, -1.0, 0.0, 0.0])
    assert a @ e == 0.0
    t = 0.5
    q = t * a + rho * e
    r = np.array([1.0, 0.3, -0.2, 0.5])
    G = np.array([[a @ a, a @ q], [a @ q, q @ q]])
    x = np.array([r @ a, r @ q])
    y32 = oracle_lib.solve_ls2x2(G, x, rank_eps=np.sqrt(np.finfo(np.float32).eps))
    G1 = np.array([[a @ a, t * (a @ a)], [t * (a @ a), t * t * (a @ a)]])
    x1 = np.array([r @ a, t * (r @ a)])
    np.testing.assert_allclose(y32, scaled_min_norm(G1, x1), rtol=10 * rho)
    # the rank-1 combination moves along a alone: y32[0] a + y32[1] q ~ (r.a / a.a) a
    np.testing.assert_allclose(y32[0] * a + y32[1] * q, (r @ a) / (a @ a) * a, rtol=10 * rho, atol=1e-12)
    y64 = oracle_lib.solve_ls2x2(G, x)
    np.testing.assert_allclose(y64[1], (r @ e) / (rho * (e @ e)), rtol=0.05)
    assert abs(y64[1]) > 1e3 * abs(y32[1])
    # the default threshold is the reference's: same result as rank_eps = DBL_EPSILON
    np.testing.assert_array_equal(y64, oracle_lib.solve_ls2x2(G, x, rank_eps=np.finfo(np.float64).eps))


@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_ls2x2_nonfinite_gives_zero(oracle_lib, bad):
    np.testing.assert_array_equal(oracle_lib.solve_ls2x2([[1.0, 0.0], [0.0, bad]], [1.0, 1.0]), [0.0, 0.0])
    np.testing.assert_array_equal(oracle_lib.solve_ls2x2([[1.0, 0.0], [0.0, 1.0]], [bad, 1.0]), [0.0, 0.0])


def test_ls2x2_zero_matrix(oracle_lib):
    np.testing.assert_array_equal(oracle_lib.solve_ls2x2(np.zeros((2, 2)), [1.0, 2.0]), [0.0, 0.0])
