"""Per-column top-k dropping (spai_column_topk, the X2 extension) against the
oracle's brute-force-checked selection (tests/test_oracle_extensions.py): the
same kept entries and values, with many exact ties in |v| (values rounded to
one decimal, both signs), the diagonal always kept, columns shorter than k
kept whole, and columns on both sides of the register path (<= 1024 entries)
and the re-reading path (longer)."""
import numpy as np
import pytest
from scipy import sparse

pytestmark = pytest.mark.gpu


def _tied(rows, cols, per_col, seed):
    rng = np.random.default_rng(seed)
    r = np.concatenate([rng.choice(rows, size=min(rows, k), replace=False) for k in per_col])
    c = np.repeat(np.arange(cols), [min(rows, k) for k in per_col])
    X = sparse.csc_matrix((np.round(rng.standard_normal(r.size), 1), (r, c)), shape=(rows, cols))
    X = (X + sparse.identity(max(rows, cols), format="csc")[:rows, :cols]).tocsc()
    X.sum_duplicates()
    X.sort_indices()
    return X


def _check(gpu_lib, oracle_lib, X, k):
    G = gpu_lib.column_topk(X, k)
    O = oracle_lib.column_topk(X, k)
    np.testing.assert_array_equal(G.indptr, O.indptr)
    np.testing.assert_array_equal(G.indices, O.indices)
    np.testing.assert_array_equal(G.data, O.data)


@pytest.mark.parametrize("k", [1, 3, 16])
def test_topk_ties_register_path(gpu_lib, oracle_lib, k):
    rng = np.random.default_rng(20 + k)
    X = _tied(3000, 600, rng.integers(0, 400, size=600), seed=k)
    print(f"case: topk ties register path k={k}", flush=True)
    _check(gpu_lib, oracle_lib, X, k)


@pytest.mark.parametrize("k", [2, 40])
def test_topk_ties_long_columns(gpu_lib, oracle_lib, k):
    rng = np.random.default_rng(30 + k)
    X = _tied(6000, 80, rng.integers(900, 3000, size=80), seed=100 + k)
    print(f"case: topk ties long columns k={k}", flush=True)
    _check(gpu_lib, oracle_lib, X, k)


def test_topk_all_equal_magnitudes(gpu_lib, oracle_lib):
    # every off-diagonal |v| equal: the k smallest rows are kept
    X = _tied(500, 300, np.full(300, 60), seed=7)
    X.data = np.where(np.arange(X.nnz) % 2 == 0, 1.5, -1.5)
    print("case: topk all equal magnitudes", flush=True)
    _check(gpu_lib, oracle_lib, X, 5)
