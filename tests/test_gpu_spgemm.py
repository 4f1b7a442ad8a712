"""The device SpGEMM of the reference-semantics path (spai_spgemm) against the
oracle's conservative product (src/common/state_manager.cpp:47-66, Eigen's
conservative sparse product; src/common/dropping.cpp:97-104 for the residual
form): the same pattern, and the same values bit for bit, since both sum the
contributions of each row in k-ascending order.

The cases reach each tier: short columns (expansion of at most 64 rows from
at most 16 columns of X: one 16-lane group sorts in registers, in one fused
pass, or in two passes past 2,097,152 such columns), the LDS bitmap (a
column's rows span fewer than 131,072 rows, at any expansion), the LDS sort
(wider span, expansion of at most 4096 rows) and the stamp array (wider span,
longer expansion).  The numeric phase holds columns of up to 1024 rows in the
LDS of a 4-wavefront block, up to 4096 in that of a 1-wavefront block, and
accumulates longer ones in global memory."""
import numpy as np
import pytest
from scipy import sparse

from parallellocallyoptimalapproximatesparsematrixinverses_amd import problems

pytestmark = pytest.mark.gpu


def _random(rows, cols, per_col, seed):
    rng = np.random.default_rng(seed)
    r = np.concatenate([rng.choice(rows, size=per_col, replace=False) for _ in range(cols)])
    c = np.repeat(np.arange(cols), per_col)
    v = rng.standard_normal(r.size)
    M = sparse.csc_matrix((v, (r, c)), shape=(rows, cols))
    M.sum_duplicates()
    M.sort_indices()
    return M


def _check(gpu_lib, oracle_lib, X, Y, residual=False, prune=False):
    C, secs = gpu_lib.spgemm(X, Y, residual=residual, prune_zero=prune, repeats=2)
    R = oracle_lib.product(X, Y, residual=residual, prune_zero=prune)
    np.testing.assert_array_equal(C.indptr, R.indptr)
    np.testing.assert_array_equal(C.indices, R.indices)
    np.testing.assert_array_equal(C.data, R.data)
    assert secs > 0.0


@pytest.mark.parametrize("residual,prune", [(False, False), (True, False), (True, True), (False, True)])
def test_spgemm_stencil_bitmap_tier(gpu_lib, oracle_lib, residual, prune):
    A = problems.diffusion3d(16, seed=1)
    print(f"case: spgemm diffusion16 residual={residual} prune={prune}", flush=True)
    _check(gpu_lib, oracle_lib, A, A, residual, prune)


def test_spgemm_random_sort_tier(gpu_lib, oracle_lib):
    # rows spread over 200,000: the span exceeds the bitmap, expansions stay small
    X = _random(200_000, 400, 40, seed=5)
    Y = _random(400, 300, 12, seed=6)
    print("case: spgemm random 200000x400 * 400x300", flush=True)
    _check(gpu_lib, oracle_lib, X, Y)


def test_spgemm_long_expansion_bitmap_tier(gpu_lib, oracle_lib):
    # 16 columns of 600 rows each per output column: expansion 9,600 > 4096
    # within 20,000 rows (the bitmap); about 7,500 rows per output column
    # (the numeric phase's global path)
    X = _random(20_000, 64, 600, seed=7)
    Y = _random(64, 40, 16, seed=8)
    print("case: spgemm random long expansion, bitmap", flush=True)
    _check(gpu_lib, oracle_lib, X, Y)


def test_spgemm_long_expansion_stamp_tier(gpu_lib, oracle_lib):
    # expansion 9,600 spread over 300,000 rows: the stamp array
    X = _random(300_000, 64, 600, seed=7)
    Y = _random(64, 40, 16, seed=8)
    print("case: spgemm random long expansion, stamp", flush=True)
    _check(gpu_lib, oracle_lib, X, Y)


@pytest.mark.parametrize("residual", [False, True])
def test_spgemm_wide_columns_in_lds(gpu_lib, oracle_lib, residual):
    # expansion 4,800 > 4096 within 5,000 rows: the bitmap, and about 3,100
    # rows per output column (the one-wavefront LDS instance of the numeric
    # phase)
    X = _random(5000, 5000, 300, seed=21)
    Y = _random(5000, 5000 if residual else 300, 16, seed=22)
    print(f"case: spgemm wide columns in LDS residual={residual}", flush=True)
    _check(gpu_lib, oracle_lib, X, Y, residual=residual)


def test_spgemm_wide_columns_with_stashed_lds_tier(gpu_lib, oracle_lib):
    # over 4096 LDS-tier columns (their rows stashed by the count pass) mixed
    # with columns of expansion over 4096 in the bitmap (rows written by the
    # fill pass)
    rng = np.random.default_rng(23)
    n = 4800
    X = _random(5000, 5000, 20, seed=24)
    per = rng.integers(17, 41, size=n)
    per[::97] = 260  # expansion 5,200
    r = np.concatenate([rng.choice(5000, size=k, replace=False) for k in per])
    c = np.repeat(np.arange(n), per)
    Y = sparse.csc_matrix((rng.standard_normal(r.size), (r, c)), shape=(5000, n))
    Y.sort_indices()
    print("case: spgemm wide columns with stashed LDS tier", flush=True)
    _check(gpu_lib, oracle_lib, X, Y)


@pytest.mark.parametrize("residual", [False, True])
def test_spgemm_mixed_tiers(gpu_lib, oracle_lib, residual):
    # Y's columns hold 1..40 entries, so short and long columns interleave
    rng = np.random.default_rng(11)
    n = 3000
    X = _random(n, n, 5, seed=12)
    per = rng.integers(1, 41, size=n)
    r = np.concatenate([rng.choice(n, size=k, replace=False) for k in per])
    c = np.repeat(np.arange(n), per)
    Y = sparse.csc_matrix((rng.standard_normal(r.size), (r, c)), shape=(n, n))
    Y.sort_indices()
    print(f"case: spgemm mixed tiers residual={residual}", flush=True)
    _check(gpu_lib, oracle_lib, X, Y, residual=residual)


@pytest.mark.parametrize("rows,residual", [(5000, False), (5000, True), (300_000, False)])
def test_spgemm_many_lds_tier_columns_stashed(gpu_lib, oracle_lib, rows, residual):
    # over 4096 columns in the LDS tiers: their count pass stashes the rows and
    # the fill pass copies them; 5000 rows use the bitmap, 300,000 the sort
    rng = np.random.default_rng(17)
    n = 5000 if residual else 4500
    X = _random(rows, 5000, 20, seed=18)
    per = rng.integers(17, 41, size=n)
    r = np.concatenate([rng.choice(5000, size=k, replace=False) for k in per])
    c = np.repeat(np.arange(n), per)
    Y = sparse.csc_matrix((rng.standard_normal(r.size), (r, c)), shape=(5000, n))
    Y.sort_indices()
    print(f"case: spgemm stashed LDS tiers rows={rows} residual={residual}", flush=True)
    _check(gpu_lib, oracle_lib, X, Y, residual=residual)


def test_spgemm_many_short_columns_two_pass(gpu_lib, oracle_lib):
    # 2,200,000 short columns: past the one-pass stash limit
    n = 2_200_000
    j = np.arange(n)
    rng = np.random.default_rng(13)

    def two_per_column(mult, add):
        rows = np.concatenate([j, (j * mult + add) % n])  # distinct: mult*j + add - j is odd, n even
        cols = np.concatenate([j, j])
        M = sparse.csc_matrix((rng.standard_normal(2 * n), (rows, cols)), shape=(n, n))
        M.sort_indices()
        return M

    X = two_per_column(7, 3)
    Y = two_per_column(13, 5)
    print("case: spgemm 2.2M short columns residual", flush=True)
    _check(gpu_lib, oracle_lib, X, Y, residual=True)


def test_spgemm_empty_columns_and_cancellation(gpu_lib, oracle_lib):
    # an empty column of Y, and two terms that cancel exactly (pruned when asked)
    X = sparse.csc_matrix(np.array([[1.0, 1.0, 0.0], [2.0, -2.0, 0.0], [0.0, 0.0, 3.0]]))
    Y = sparse.csc_matrix(np.array([[1.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]]))
    for prune in (False, True):
        _check(gpu_lib, oracle_lib, X, Y, prune=prune)
    C, _ = gpu_lib.spgemm(X, Y, prune_zero=True)
    assert C.nnz == 2  # (0,0) = 2 and (2,2) = 3; (1,0) = 0 is pruned


def test_spgemm_dimension_errors(gpu_lib):
    X = sparse.csc_matrix(np.eye(3))
    Y = sparse.csc_matrix(np.eye(4))
    with pytest.raises(RuntimeError, match="dimensions do not match"):
        gpu_lib.spgemm(X, Y)
