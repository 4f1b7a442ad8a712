"""Shared comparisons for the parity tests."""
import contextlib
import os

import numpy as np
from scipy import sparse


def tri(n):
    """tests/test.cpp:13-28 fixture: tridiag(-1, 2, -1)."""
    return sparse.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1], format="csc")


def rel_fro(a, b):
    d = sparse.linalg.norm(sparse.csc_matrix(a) - sparse.csc_matrix(b)) if sparse.issparse(a) else np.linalg.norm(a - b)
    nb = sparse.linalg.norm(b) if sparse.issparse(b) else np.linalg.norm(b)
    return d / nb if nb > 0 else d


def assert_identical(g, o):
    """Bit-for-bit equality of two results (history and M)."""
    assert g["iterations"] == o["iterations"] and g["converged"] == o["converged"]
    for f in ("iteration", "residual_norm", "density_m", "density_direction"):
        np.testing.assert_array_equal(g["history"][f], o["history"][f], err_msg=f)
    for f in ("indptr", "indices", "data"):
        np.testing.assert_array_equal(getattr(g["M"], f), getattr(o["M"], f), err_msg=f)


def assert_same_result(g, o, hist_rtol=1e-10, m_rtol=1e-12, same_pattern=True, direction_pattern=True):
    """Device result g against oracle result o (dicts from _lib.run / oracle.run)."""
    assert g["iterations"] == o["iterations"], (g["iterations"], o["iterations"])
    assert g["converged"] == o["converged"]
    hg, ho = g["history"], o["history"]
    assert len(hg) == len(ho)
    np.testing.assert_array_equal(hg["iteration"], ho["iteration"])
    np.testing.assert_allclose(hg["residual_norm"], ho["residual_norm"], rtol=hist_rtol, atol=1e-300)
    if same_pattern:
        np.testing.assert_array_equal(hg["density_m"], ho["density_m"])
        if direction_pattern:
            np.testing.assert_array_equal(hg["density_direction"], ho["density_direction"])
        Mg, Mo = g["M"], o["M"]
        np.testing.assert_array_equal(Mg.indptr, Mo.indptr)
        np.testing.assert_array_equal(Mg.indices, Mo.indices)
    assert rel_fro(g["M"], o["M"]) <= m_rtol, rel_fro(g["M"], o["M"])


def dense_loss_reference(A, M):
    """numpy FP64 f = ||I - A M||_F^2 and its gradients sampled on the patterns
    (df/dM = -2 A^T R on pattern(M), df/dA = -2 R M^T on pattern(A)), CSC order."""
    A = sparse.csc_matrix(A)
    M = sparse.csc_matrix(M)
    A.sort_indices()
    M.sort_indices()
    Ad, Md = A.toarray(), M.toarray()
    R = np.eye(A.shape[0]) - Ad @ Md
    GM = -2.0 * (Ad.T @ R)
    GA = -2.0 * (R @ Md.T)
    gm = np.array([GM[M.indices[p], j] for j in range(M.shape[1]) for p in range(M.indptr[j], M.indptr[j + 1])])
    ga = np.array([GA[A.indices[p], j] for j in range(A.shape[1]) for p in range(A.indptr[j], A.indptr[j + 1])])
    return float((R * R).sum()), gm, ga


@contextlib.contextmanager
def env_set(name, value):
    """Environment variable name = value inside the block (restored after)."""
    old = os.environ.get(name)
    os.environ[name] = value
    try:
        yield
    finally:
        if old is None:
            del os.environ[name]
        else:
            os.environ[name] = old
