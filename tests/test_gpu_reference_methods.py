"""Device builders (reference semantics) against the oracle, through the C ABI.

Fixtures follow tests/test.cpp:107-179 (n=12 tridiagonal, M0 = Pr = I) plus
random SPD matrices and the reference's sample matrices.  Two bars:

* reference-named entry points (tree reductions): same iteration counts and
  stop flags, identical M patterns and densities, residual histories within
  1e-10 and M within 1e-12 (relative Frobenius).  The direction's pattern may
  differ only at entries within eps*||P|| of the prune threshold (CG with
  dropping), so it is compared on the batch-ordered runs.
* batch_ordered_reductions = 1 (reductions in the reference's order): the
  device result equals the oracle's bit for bit, and reproduces the
  reference's committed histories.
"""
import numpy as np
import pytest
from scipy import sparse

import golden_data as gd
from helpers import assert_identical, assert_same_result, tri

pytestmark = pytest.mark.gpu

METHODS = ["cg", "mr", "lomr", "io_mr", "io_lomr"]


def both(gpu_lib, oracle_lib, method, *args, batch_ordered=False, **kw):
    print(f"case: {method} {args[3:] if len(args) > 3 else ''} {kw} ordered={batch_ordered}", flush=True)
    if batch_ordered:
        g = gpu_lib.run(method, *args, batch_ordered=True, **kw)
    else:
        g = gpu_lib.run(method, *args, plain_entry=True, **kw)
    o = oracle_lib.run(method, *args, **kw)
    return g, o


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("threads", [1, 2])
def test_parallelization_fixture(gpu_lib, oracle_lib, method, threads):
    A, I = tri(12), sparse.identity(12, format="csc")
    g, o = both(gpu_lib, oracle_lib, method, A, I, I, 8, 1e-14, 1.0, True, threads)
    assert_same_result(g, o, direction_pattern=False)
    assert g["history"]["residual_norm"][-1] < g["history"]["residual_norm"][0]
    g, o = both(gpu_lib, oracle_lib, method, A, I, I, 8, 1e-14, 1.0, True, threads, batch_ordered=True)
    assert_identical(g, o)


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


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("dropping,density", [(False, 1.0), (True, 0.05)])
def test_random_spd_batch_ordered_identical(gpu_lib, oracle_lib, method, dropping, density):
    A = random_spd(150, 3, 7)
    Pr = sparse.diags(1.0 / A.diagonal(), format="csc")
    g, o = both(gpu_lib, oracle_lib, method, A, None, Pr, 4, 1e-12, density, dropping, 8, batch_ordered=True)
    assert_identical(g, o)


GOLDEN = [("tri100eigs4k", "no_dropping", "1.000", 8), ("tri100eigs4k", "dropping", "0.030", 8),
          ("bcsstk21", "no_dropping", "1.000", 4), ("bcsstk21", "dropping", "0.030", 4),
          ("msc04515", "dropping", "0.030", 6)]


@pytest.mark.parametrize("name,mode,density,iters", GOLDEN)
@pytest.mark.parametrize("variant", range(7))
def test_device_reproduces_reference_history(gpu_lib, oracle_lib, name, mode, density, iters, variant):
    """Device runs (batch-ordered, 256 batches) against the reference's committed
    convergence tables and against the oracle, bit for bit."""
    method, inner = gd.VARIANTS[variant]
    if method.startswith("io") and name == "bcsstk21":
        iters = 2
    A = gd.matrix(name)
    M0, Pr = gd.harness_inputs(A)
    g, o = both(gpu_lib, oracle_lib, method, A, M0, Pr, iters, 1e-9, float(density), mode == "dropping", 256,
                inner_iterations=inner, batch_ordered=True)
    assert_identical(g, o)
    ref = gd.convergence(name, mode, density)[:, variant]
    h = g["history"]["residual_norm"]
    ok = ~np.isnan(ref[:len(h)])
    np.testing.assert_allclose(h[ok], ref[:len(h)][ok], rtol=1e-12, atol=0)
