"""tests/test.cpp:107-179 invariants, re-expressed on the oracle."""
import numpy as np
import pytest
from scipy import sparse

from helpers import tri

METHODS = ["cg", "mr", "lomr", "io_mr", "io_lomr"]
A12 = tri(12)
I12 = sparse.identity(12, format="csc")


@pytest.mark.parametrize("method", METHODS)
def test_parallelization(oracle_lib, method):
    # tests/test.cpp:74-81,120-127
    r1 = oracle_lib.run(method, A12, I12, I12, 8, 1e-14, 1.0, True, 1)
    r2 = oracle_lib.run(method, A12, I12, I12, 8, 1e-14, 1.0, True, 2)
    for r in (r1, r2):
        h = r["history"]
        assert len(h) == r["iterations"] + 1 and h["iteration"][0] == 0 and h["iteration"][-1] == r["iterations"]
        assert h["residual_norm"][-1] < np.sqrt(34.0)
    assert r1["iterations"] == r2["iterations"]
    assert sparse.linalg.norm(r1["M"] - r2["M"]) < 1e-10
    assert abs(r1["history"]["residual_norm"][-1] - r2["history"]["residual_norm"][-1]) < 1e-10


def test_initial_residual(oracle_lib):
    # tests/test.cpp:98-103,112: ||I - A||_F = sqrt(34)
    r = oracle_lib.run("mr", A12, I12, I12, 1, 0.0, 1.0, True, 1)
    assert r["history"]["residual_norm"][0] == pytest.approx(np.sqrt(34.0), rel=1e-15)


def test_absolute_stop(oracle_lib):
    # tests/test.cpp:129-138
    tol = 0.9 * np.sqrt(34.0)
    r = oracle_lib.run("mr", A12, I12, I12, 8, tol, 1.0, True, 2)
    assert r["converged"] and r["iterations"] == 1 and len(r["history"]) == 2
    assert r["history"]["residual_norm"][-1] < tol


@pytest.mark.parametrize("method", METHODS)
def test_dropping_budget(oracle_lib, method):
    # tests/test.cpp:140-151: density 0.12 -> budget 17, exactly 3 iterations, 1 vs 4 batches
    r1 = oracle_lib.run(method, A12, I12, I12, 3, 0.0, 0.12, True, 1)
    r4 = oracle_lib.run(method, A12, I12, I12, 3, 0.0, 0.12, True, 4)
    for r in (r1, r4):
        assert r["M"].nnz <= 17 and r["iterations"] == 3
    assert sparse.linalg.norm(r1["M"] - r4["M"]) < 1e-10


@pytest.mark.parametrize("method", METHODS)
def test_density_stop(oracle_lib, method):
    # tests/test.cpp:153-165
    r = oracle_lib.run(method, A12, I12, I12, 8, 0.0, 0.2, False, 1)
    assert not r["converged"] and 0 < r["iterations"] < 8
    assert r["history"]["density_m"][-1] >= 0.2


def test_reported_matches_true_residual(oracle_lib):
    # tests/test.cpp:167-177
    r = oracle_lib.run("io_lomr", A12, I12, I12, 20, 1e-20, 1.0, True, 2)
    assert r["iterations"] == 20 and not r["converged"]
    true = np.linalg.norm((sparse.identity(12) - A12 @ r["M"]).toarray())
    assert abs(r["history"]["residual_norm"][-1] - true) < 1e-9


@pytest.mark.parametrize("kwargs,msg", [
    (dict(max_iterations=0), "max_iterations must be positive"),
    (dict(tolerance=-1.0), "tolerance must be finite and non-negative"),
    (dict(max_density=0.0), "max_density must be finite and in the interval (0, 1]"),
    (dict(max_density=0.01), "max_density is too small to preserve the diagonal of M"),
    (dict(num_threads=0), "Invalid threads count value"),
])
def test_validation_messages(oracle_lib, kwargs, msg):
    # src/common/methods_common.cpp:12-48
    base = dict(max_iterations=3, tolerance=0.0, max_density=0.5, enable_dropping=True, num_threads=1)
    base.update(kwargs)
    with pytest.raises(RuntimeError, match=msg.replace("(", r"\(").replace(")", r"\)").replace("]", r"\]")):
        oracle_lib.run("mr", A12, I12, I12, **base)


def test_invalid_diagonal(oracle_lib):
    A = tri(12).tolil()
    A[3, 3] = -1.0
    with pytest.raises(RuntimeError, match="invalid diagonal"):
        oracle_lib.run("mr", A.tocsc(), None, None, 2, 0.0, 1.0, False, 1)
