"""Sample-dataset runner: the reference's construction + backward-error study
on one matrix, with every builder and PCG solve on the device.

Mirrors scripts/common.py:
  * load_matrix_market        _load_sparse_matrix (common.py:396-400)
  * harness_inputs            MatrixData.load_data (common.py:380-385, 403-408)
  * construction_row          the construction_status.csv row (common.py:240-263)
  * run_backward_error        the backward_error.csv histories and the
                              solver_status.csv rows (common.py:146-167)
  * run_study                 both, over the seven method variants in the
                              column order of the reference's tables
                              (scripts/utils/methods.py:17-37)

Host-side bookkeeping (canonicalisation, ||I - A M||_F of the finished M) uses
scipy as the reference's scripts do; the builders and the PCG run through
libspai_hip.so.
"""
from dataclasses import dataclass, field

import numpy as np
from scipy import sparse
from scipy.io import mmread
from scipy.sparse.linalg import norm as sparse_norm

from . import _lib, harness

# (label, method, inner iterations) in the column order of the reference's
# convergence and backward-error tables.
VARIANTS = [
    ("Global CG", "cg", 2),
    ("Global MR", "mr", 2),
    ("Global LOMR", "lomr", 2),
    ("Inner-outer - 2 inner iter MR", "io_mr", 2),
    ("Inner-outer - 2 inner iter LOMR", "io_lomr", 2),
    ("Inner-outer - 5 inner iter MR", "io_mr", 5),
    ("Inner-outer - 5 inner iter LOMR", "io_lomr", 5),
]


def load_matrix_market(path) -> sparse.csc_matrix:
    """CSC float64 with duplicates summed, explicit zeros removed and row
    indices sorted (scripts/common.py:396-400)."""
    A = mmread(str(path)).tocsc().astype(np.float64)
    A.sum_duplicates()
    A.eliminate_zeros()
    A.sort_indices()
    return A


def harness_inputs(A):
    """M0 = I and Pr = diag(A)^-1; A needs a finite positive diagonal
    (scripts/common.py:380-385, 403-408)."""
    diag = A.diagonal()
    if not np.isfinite(diag).all() or np.any(diag <= 0.0):
        raise ValueError("A must have a finite positive diagonal.")
    n = A.shape[0]
    return sparse.eye(n, format="csc"), sparse.diags(1.0 / diag, format="csc")


def construction_row(label, A, M, result, num_threads, tolerance):
    """One construction_status.csv row (scripts/common.py:251-263)."""
    n = A.shape[0]
    actual = sparse_norm(sparse.eye(n, format="csc") - A @ M)
    return {
        "method": label,
        "requested_threads": num_threads,
        "iterations": int(result["iterations"]),
        "reported_residual_norm": float(result["history"]["residual_norm"][-1]),
        "actual_residual_norm": float(actual),
        "construction_backward_error": float(actual / (sparse_norm(A) * sparse_norm(M) + np.sqrt(n))),
        "density": M.nnz / n ** 2,
        "converged": bool(actual < tolerance),
    }


def run_backward_error(A, preconditioners, max_iterations=150, seed=1, realizations=1, num_threads=1):
    """PCG with each (label, M) on the same right-hand sides: per realization
    the list of (label, history), and the solver_status.csv rows
    (scripts/common.py:146-167)."""
    system = harness.PreparedPCGSystem(A, num_threads=num_threads)
    limit = max_iterations or A.shape[0]
    histories = [[] for _ in range(realizations)]
    rows = []
    for label, M in preconditioners:
        solver = system.prepare_preconditioner(M)
        try:
            rhs = harness.random_rhs_realizations(system.A, seed=seed, realizations=realizations,
                                                  num_threads=system.num_threads)
            for realization, b in enumerate(rhs):
                ev = solver.calculate(max_iterations, b=b)
                rows.append({
                    "method": label,
                    "realization": realization + 1,
                    "solver": ev.solver,
                    "active_threads": system.active_threads,
                    "history_limit": limit,
                    "solver_updates": ev.history.size - 1,
                    "final_relative_residual": float(ev.history[-1]),
                    "converged": ev.converged,
                    "status": ev.status,
                })
                histories[realization].append((label, ev.history))
        finally:
            solver.close()
    return histories, rows


@dataclass
class Study:
    convergence: list = field(default_factory=list)    # (label, ||R_i||_F history)
    construction: list = field(default_factory=list)   # construction_status rows
    backward_error: list = field(default_factory=list)  # per realization: (label, PCG history)
    solver_status: list = field(default_factory=list)   # solver_status rows
    preconditioners: list = field(default_factory=list)  # (label, M)


def run_study(A, max_iterations, max_density, enable_dropping, tolerance=1e-9, num_threads=256,
              variants=None, batch_ordered=True, backward_iterations=150, seed=1, realizations=1):
    """The reference's sample-dataset study on A (scripts/common.py:225-263
    and 140-167) with the builders and PCG on the device.  batch_ordered=True
    sums the trace products in the reference's column-batch order, which makes
    the histories bit-identical to the reference's at the same thread count."""
    A = load_csc(A)
    M0, Pr = harness_inputs(A)
    study = Study()
    for label, method, inner in (variants or VARIANTS):
        res = _lib.run(method, A, M0, Pr, max_iterations, tolerance, max_density, enable_dropping, num_threads,
                       inner_iterations=inner, batch_ordered=batch_ordered)
        M = res["M"]
        study.convergence.append((label, res["history"]["residual_norm"]))
        study.construction.append(construction_row(label, A, M, res, num_threads, tolerance))
        study.preconditioners.append((label, M))
    study.backward_error, study.solver_status = run_backward_error(
        A, study.preconditioners, max_iterations=backward_iterations, seed=seed, realizations=realizations,
        num_threads=num_threads)
    return study


def load_csc(A) -> sparse.csc_matrix:
    A = sparse.csc_matrix(A, dtype=np.float64)
    A.sort_indices()
    return A
