"""Loaders of the golden fixtures in tests/golden/."""
import csv
from pathlib import Path

import numpy as np
from scipy import sparse

GOLDEN = Path(__file__).resolve().parent / "golden"

# Column order of the reference's convergence tables (scripts/utils/methods.py:17-37).
VARIANTS = [("cg", 2), ("mr", 2), ("lomr", 2), ("io_mr", 2), ("io_lomr", 2), ("io_mr", 5), ("io_lomr", 5)]


def matrix(name):
    """The canonical CSC matrix (sorted indices).  Fixtures flagged "lower" hold
    the lower triangle of an exactly symmetric matrix; the full matrix is
    L + strict_lower(L)^T, whose two terms share no position."""
    z = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    A = sparse.csc_matrix((z["data"], z["indices"].astype(np.int64), z["indptr"].astype(np.int64)),
                          shape=tuple(z["shape"]))
    if "lower" in z.files:
        A = (A + sparse.tril(A, k=-1).T).tocsc()
        A.sort_indices()
    return A


def convergence(name, mode, density):
    """Array [iteration, variant] of ||R_i||_F (NaN where a run had stopped)."""
    with open(GOLDEN / f"{name}-{mode}-d{density}-convergence.csv") as fh:
        rows = list(csv.reader(fh))
    return np.array([[float(x) if x else np.nan for x in r[1:]] for r in rows[1:]])


def harness_inputs(A):
    """M0 = I and Pr = diag(A)^-1 as scripts/common.py:380-409 sets them."""
    n = A.shape[0]
    return sparse.identity(n, format="csc"), sparse.diags(1.0 / A.diagonal(), format="csc")
