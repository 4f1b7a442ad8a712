"""Device builders (reference semantics) against the oracle, through the C ABI.

Fixtures follow tests/test.cpp:107-179 (n=12 tridiagonal, M0 = Pr = I) plus
random SPD matrices; every comparison is device vs the CPU restatement on the
same inputs: same iteration counts and stop flags, identical M patterns and
densities, residual histories within 1e-10 and M within 1e-12 (relative
Frobenius).
"""
import numpy as np
import pytest
from scipy import sparse

from helpers import assert_same_result, tri

pytestmark = pytest.mark.gpu

METHODS = ["cg", "mr", "lomr", "io_mr", "io_lomr"]


def both(gpu_lib, oracle_lib, method, *args, **kw):
    print(f"case: {method} {args[3:] if len(args) > 3 else ''} {kw}", flush=True)
    g = gpu_lib.run(method, *args, plain_entry=True, **kw)
    o = oracle_lib.run(method, *args, **kw)
    return g, o


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("threads", [1, 2])
def test_parallelization_fixture(gpu_lib, oracle_lib, method, threads):
    A, I = tri(12), sparse.identity(12, format="csc")
    g, o = both(gpu_lib, oracle_lib, method, A, I, I, 8, 1e-14, 1.0, True, threads)
    assert_same_result(g, o)
    assert g["history"]["residual_norm"][-1] < g["history"]["residual_norm"][0]


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("threads", [1, 4])
def test_dropping_fixture(gpu_lib, oracle_lib, method, threads):
    A, I = tri(12), sparse.identity(12, format="csc")
    g, o = both(gpu_lib, oracle_lib, method, A, I, I, 3, 0.0, 0.12, True, threads)
    assert_same_result(g, o)
    assert g["M"].nnz <= 17 and g["iterations"] == 3


@pytest.mark.parametrize("method", METHODS)
def test_no_dropping_density_stop(gpu_lib, oracle_lib, method):
    A, I = tri(12), sparse.identity(12, format="csc")
    g, o = both(gpu_lib, oracle_lib, method, A, I, I, 8, 0.0, 0.2, False, 1)
    assert_same_result(g, o)
    assert not g["converged"] and 0 < g["iterations"] < 8
    assert g["history"]["density_m"][-1] >= 0.2


def test_absolute_stop(gpu_lib, oracle_lib):
    A, I = tri(12), sparse.identity(12, format="csc")
    tol = 0.9 * np.sqrt(34.0)
    g, o = both(gpu_lib, oracle_lib, "mr", A, I, I, 8, tol, 1.0, True, 2)
    assert_same_result(g, o)
    assert g["converged"] and g["iterations"] == 1


def test_reported_vs_true_residual(gpu_lib, oracle_lib):
    A, I = tri(12), sparse.identity(12, format="csc")
    g, o = both(gpu_lib, oracle_lib, "io_lomr", A, I, I, 20, 1e-20, 1.0, True, 2)
    assert_same_result(g, o, hist_rtol=1e-8, m_rtol=1e-10)
    true = np.linalg.norm((sparse.identity(12) - A @ g["M"]).toarray())
    assert abs(g["history"]["residual_norm"][-1] - true) < 1e-9


def random_spd(n, per_col, seed):
    rng = np.random.default_rng(seed)
    rows = rng.integers(0, n, size=n * per_col)
    cols = np.repeat(np.arange(n), per_col)
    B = sparse.csc_matrix((rng.standard_normal(n * per_col), (rows, cols)), shape=(n, n))
    A = (B.T @ B + sparse.identity(n) * 0.5).tocsc()
    A.sort_indices()
    return A


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("dropping,density", [(False, 1.0), (True, 0.05)])
def test_random_spd(gpu_lib, oracle_lib, method, dropping, density):
    A = random_spd(150, 3, 7)
    Pr = sparse.diags(1.0 / A.diagonal(), format="csc")
    g, o = both(gpu_lib, oracle_lib, method, A, None, Pr, 4, 1e-12, density, dropping, 8)
    assert_same_result(g, o, hist_rtol=1e-9, m_rtol=1e-10)
