"""Host logic of the sample-dataset runner (no device calls)."""
import numpy as np
import pytest
from scipy import sparse
from scipy.io import mmwrite

from parallellocallyoptimalapproximatesparsematrixinverses_amd import datasets


def test_load_matrix_market_canonicalises(tmp_path):
    # duplicates summed, explicit zeros dropped, rows sorted (scripts/common.py:396-400)
    rows = np.array([2, 0, 0, 1, 2, 1])
    cols = np.array([0, 0, 0, 1, 2, 2])
    vals = np.array([1.0, 2.0, 3.0, 0.0, 4.0, 5.0])
    A = sparse.coo_matrix((vals, (rows, cols)), shape=(3, 3))
    path = tmp_path / "a.mtx"
    mmwrite(str(path), A)
    B = datasets.load_matrix_market(path)
    assert isinstance(B, sparse.csc_matrix) and B.dtype == np.float64
    assert B.has_sorted_indices
    assert B.nnz == 4
    np.testing.assert_array_equal(B.toarray(), np.array([[5.0, 0, 0], [0, 0, 5.0], [1.0, 0, 4.0]]))


def test_harness_inputs():
    A = sparse.csc_matrix(np.array([[4.0, 1.0], [1.0, 2.0]]))
    M0, Pr = datasets.harness_inputs(A)
    np.testing.assert_array_equal(M0.toarray(), np.eye(2))
    np.testing.assert_array_equal(Pr.toarray(), np.diag([0.25, 0.5]))
    with pytest.raises(ValueError, match="finite positive diagonal"):
        datasets.harness_inputs(sparse.csc_matrix(np.array([[1.0, 0.0], [0.0, 0.0]])))


def test_construction_row_formulas():
    A = sparse.csc_matrix(np.array([[2.0, 1.0], [1.0, 3.0]]))
    M = sparse.csc_matrix(np.array([[0.5, 0.0], [0.0, 1.0 / 3.0]]))
    res = {"iterations": 7, "history": np.array([(0, 3.0, 0.5, 0.0), (7, 0.25, 0.5, 0.0)],
                                                  dtype=[("iteration", "i8"), ("residual_norm", "f8"),
                                                         ("density_m", "f8"), ("density_direction", "f8")])}
    row = datasets.construction_row("Global MR", A, M, res, 256, 1e-9)
    R = np.eye(2) - A.toarray() @ M.toarray()
    actual = np.linalg.norm(R)
    assert row["method"] == "Global MR" and row["requested_threads"] == 256 and row["iterations"] == 7
    assert row["reported_residual_norm"] == 0.25
    np.testing.assert_allclose(row["actual_residual_norm"], actual, rtol=1e-15)
    np.testing.assert_allclose(row["construction_backward_error"],
                               actual / (np.linalg.norm(A.toarray()) * np.linalg.norm(M.toarray()) + np.sqrt(2)),
                               rtol=1e-15)
    assert row["density"] == 0.5 and row["converged"] is False


def test_variant_order_matches_reference_tables():
    import golden_data as gd
    header = open(gd.GOLDEN / "tri100eigs4k-no_dropping-d1.000-convergence.csv").readline().strip().split(",")
    assert header[1:] == [v[0] for v in datasets.VARIANTS]
    assert [(m, i) for _, m, i in datasets.VARIANTS] == gd.VARIANTS
