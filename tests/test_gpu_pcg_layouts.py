"""Row layouts of the device PCG's iteration kernels.

The solver reads A and M in one of three layouts: diagonals (matrices whose
entries lie on at most 32 diagonals, as grid stencils do), sliced ELLPACK
(short rows) or CSR.  Each sums a row's entries in column order, the order
of backward_error.pyx:16-34, and the library is built without contracted
multiply-adds, so the three must give the same bits: histories, statuses and
iterates are compared for exact equality, and the diagonal layout against
the oracle's PCG as the other PCG tests do (history 1e-6, counts +-1).
"""
import os

import numpy as np
import pytest
from scipy import sparse

from parallellocallyoptimalapproximatesparsematrixinverses_amd import harness, problems

pytestmark = pytest.mark.gpu


def _solve(A, M, b, limit, tol, m_fp32, layout):
    """(layouts, history, status, x) of one solve with SPAI_PCG_LAYOUT = layout."""
    old = os.environ.get("SPAI_PCG_LAYOUT")
    os.environ["SPAI_PCG_LAYOUT"] = layout
    try:
        sysm = harness._DeviceSystem(sparse.csr_matrix(A), sparse.csr_matrix(M), m_fp32=m_fp32)
    finally:
        if old is None:
            del os.environ["SPAI_PCG_LAYOUT"]
        else:
            os.environ["SPAI_PCG_LAYOUT"] = old
    try:
        hist, st, x = sysm.solve(b, limit, tol)
        return sysm.layouts(), hist, st, x
    finally:
        sysm.close()


def _spai_like(A, seed):
    """A symmetric positive definite preconditioner on pattern(A):
    diag(A)^-1 plus small values on the off-diagonal entries."""
    A = sparse.csr_matrix(A)
    M = A.copy()
    rng = np.random.default_rng(seed)
    d = A.diagonal()
    M.data = rng.uniform(-0.02, 0.02, M.nnz) / np.abs(d[M.indices])
    M = sparse.csr_matrix(0.5 * (M + M.T))
    M.setdiag(1.0 / d)
    M.sort_indices()
    return M


def _spai(oracle_lib, A):
    """The oracle's fixed-pattern MR preconditioner (10 iterations), CSR."""
    Pr = problems.diag_inverse(A)
    return oracle_lib.run("mr", A, Pr, None, 10, 0.0, 1.0, False, 1, fixed_pattern=True)["M"].tocsr()


def _with_far_pair(A):
    """A plus one symmetric pair of entries on the outermost diagonals: two
    diagonals with a single entry each (mask bits set in one row only)."""
    A = sparse.lil_matrix(A)
    n = A.shape[0]
    A[0, n - 1] = -1e-3
    A[n - 1, 0] = -1e-3
    return sparse.csc_matrix(A)


CASES = {
    "diffusion3d_20": lambda: problems.diffusion3d(20, seed=1),
    "poisson2d_40": lambda: problems.poisson2d(40),
    "stencil27_16": lambda: problems.stencil27(16, seed=4),
    "diffusion3d_12_far_pair": lambda: _with_far_pair(problems.diffusion3d(12, seed=1)),
}


@pytest.mark.parametrize("m_fp32", [False, True])
@pytest.mark.parametrize("case", sorted(CASES))
def test_layouts_give_the_same_bits(oracle_lib, case, m_fp32):
    A = CASES[case]()
    M = _spai(oracle_lib, A)
    n = A.shape[0]
    b = A @ np.random.default_rng(2).standard_normal(n)
    print(f"case: pcg layouts {case} m_fp32={m_fp32} n={n} nnz={A.nnz}", flush=True)
    runs = {lay: _solve(A, M, b, 3000, 1e-8, m_fp32, lay) for lay in ("auto", "sell", "csr")}
    assert runs["auto"][0] == ("dia", "dia"), runs["auto"][0]
    assert runs["csr"][0] == ("csr", "csr")
    _, h, st, x = runs["auto"]
    if A.nnz <= 16 * n:
        # short rows: SELL and CSR (one thread per row) sum in column order
        assert runs["sell"][0] == ("sell", "sell")
        for lay in ("sell", "csr"):
            _, h2, st2, x2 = runs[lay]
            assert st == st2
            np.testing.assert_array_equal(h, h2, err_msg=lay)
            np.testing.assert_array_equal(x, x2, err_msg=lay)
    else:
        # longer rows (27-point stencil): CSR takes 16 lanes per row and
        # adds in another order; the diagonal layout keeps column order
        assert runs["sell"][0] == ("csr", "csr")
    Mo = sparse.csr_matrix(np.float32(M.toarray()).astype(np.float64)) if m_fp32 else M
    ho, so, _ = oracle_lib.pcg(sparse.csr_matrix(A), sparse.csr_matrix(Mo), b, 3000, 1e-8)
    assert st == so == 1
    assert abs(len(h) - len(ho)) <= 1
    k = min(len(h), len(ho)) - 1
    np.testing.assert_allclose(h[:k], ho[:k], rtol=1e-6, atol=0)


def test_layout_fallbacks():
    """Matrices off the diagonal layout: many diagonals (random Gram) take
    SELL; a single far entry pair on a tiny matrix keeps the diagonals."""
    A = problems.random_gram(3000, per_row=3, seed=3, power_iterations=10)[0]
    M = _spai_like(A, 6)
    b = A @ np.random.default_rng(2).standard_normal(A.shape[0])
    lay, h, st, x = _solve(A, M, b, 500, 1e-8, False, "auto")
    assert lay == ("sell", "sell"), lay
    lay2, h2, st2, x2 = _solve(A, M, b, 500, 1e-8, False, "csr")
    assert st == st2
    np.testing.assert_array_equal(h, h2)
    np.testing.assert_array_equal(x, x2)


@pytest.mark.parametrize("m_fp32", [False, True])
def test_diagonal_layout_several_rows_per_thread(m_fp32):
    """n = 1,210,000 > the resident grid's 1,048,576 threads: each thread
    takes a second row after its first (operands of the first row loaded
    before the kernels' prologues).  60 iterations, compared bit for bit
    with the SELL layout."""
    A = problems.anisotropic2d(1100)
    M = _spai_like(A, 7)
    b = A @ np.random.default_rng(2).standard_normal(A.shape[0])
    print(f"case: pcg dia several rows per thread m_fp32={m_fp32}", flush=True)
    lay, h, st, x = _solve(A, M, b, 60, 1e-12, m_fp32, "auto")
    assert lay == ("dia", "dia")
    lay2, h2, st2, x2 = _solve(A, M, b, 60, 1e-12, m_fp32, "sell")
    assert lay2 == ("sell", "sell")
    assert st == st2 == 0 and len(h) == 60
    np.testing.assert_array_equal(h, h2)
    np.testing.assert_array_equal(x, x2)
