"""ORACLE -- TEST INFRASTRUCTURE ONLY.

ctypes front-end of oracle/_build/liboracle_spai.so, the CPU restatement of the
reference algorithms (oracle/spai_oracle.cpp).  Only tests/, the smoke check
and the benchmark's CPU-baseline leg import this module; the product never does.

Parity pins (what checks each oracle function; tests named in brackets):
  * conservative product / residual (src/common/state_manager.cpp:60,
    src/common/dropping.cpp:97-104): pattern equal to the Boolean product of
    the stored patterns (explicit zeros kept), values against exact rational
    products, no transposition, residual diagonal insertion and zero pruning
    [test_oracle_ops.py].
  * SolveLeastSquares2x2 (src/common/sparse_operations.cpp:58-73): numpy solve
    for nonsingular Grams, the minimum-norm solution in the diagonally scaled
    coordinates for rank 1, zero on non-finite input [test_oracle_ops.py].
    The exact rank threshold of Eigen's decomposition near singularity is
    "parity unpinned" (pivots within a few ulps of 2 eps * max pivot).
  * Global CG/MR/LOMR and inner-outer MR/LOMR end to end: the reference's own
    committed per-iteration ||R_i||_F histories (tests/golden/) at 256 column
    batches [test_oracle_golden.py], the tests/test.cpp invariants
    [test_oracle_invariants.py], and a dense pure-Python restatement on small
    matrices [test_oracle_dense.py].
  * PCG (methods_cython/backward_error.pyx:69-144): exact termination in
    <= n steps, the A-norm optimality of x_k over the preconditioned Krylov
    space, history[0] = 1 with a final true-residual entry, status codes
    [test_oracle_pcg.py].
  * Extensions (fixed pattern, per-column top-k): projected line-search
    optimality, two-term normal equations, projected CG, brute-force sort
    [test_oracle_extensions.py].
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np
from scipy import sparse

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle_spai.so"

METHODS = {"cg": 0, "mr": 1, "lomr": 2, "io_mr": 3, "io_lomr": 4}


class MatrixData(ctypes.Structure):
    _fields_ = [
        ("rows", ctypes.c_int64),
        ("cols", ctypes.c_int64),
        ("nnz", ctypes.c_int64),
        ("col_offsets", ctypes.POINTER(ctypes.c_int64)),
        ("row_indices", ctypes.POINTER(ctypes.c_int64)),
        ("values", ctypes.POINTER(ctypes.c_double)),
    ]


class MethodParams(ctypes.Structure):
    _fields_ = [
        ("max_iterations", ctypes.c_int64),
        ("tolerance", ctypes.c_double),
        ("max_density", ctypes.c_double),
        ("enable_dropping", ctypes.c_int64),
        ("num_threads", ctypes.c_int64),
        ("use_initial_m", ctypes.c_int64),
        ("use_preconditioner", ctypes.c_int64),
        ("inner_iterations", ctypes.c_int64),
    ]


class SingleIterationData(ctypes.Structure):
    _fields_ = [
        ("iteration", ctypes.c_int64),
        ("residual_norm", ctypes.c_double),
        ("density_m", ctypes.c_double),
        ("density_direction", ctypes.c_double),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("rows", ctypes.c_int64),
        ("cols", ctypes.c_int64),
        ("nnz", ctypes.c_int64),
        ("iterations", ctypes.c_int64),
        ("converged", ctypes.c_int64),
        ("history_size", ctypes.c_int64),
        ("col_offsets", ctypes.POINTER(ctypes.c_int64)),
        ("row_indices", ctypes.POINTER(ctypes.c_int64)),
        ("values", ctypes.POINTER(ctypes.c_double)),
        ("history", ctypes.POINTER(SingleIterationData)),
        ("error_message", ctypes.c_char_p),
    ]


class SpaiExtParams(ctypes.Structure):
    _fields_ = [
        ("fixed_pattern", ctypes.c_int64),
        ("precision", ctypes.c_int64),
        ("column_top_k", ctypes.c_int64),
        ("batch_ordered_reductions", ctypes.c_int64),
        ("reserved", ctypes.c_int64 * 4),
    ]


class CsrData(ctypes.Structure):
    _fields_ = [
        ("rows", ctypes.c_int64),
        ("cols", ctypes.c_int64),
        ("nnz", ctypes.c_int64),
        ("row_offsets", ctypes.POINTER(ctypes.c_int64)),
        ("col_indices", ctypes.POINTER(ctypes.c_int64)),
        ("values", ctypes.POINTER(ctypes.c_double)),
    ]


HISTORY_DTYPE = np.dtype([
    ("iteration", np.int64),
    ("residual_norm", np.float64),
    ("density_m", np.float64),
    ("density_direction", np.float64),
])

_lib = None


def build() -> Path:
    """Compiles the oracle library (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH), mode=ctypes.RTLD_LOCAL)
        L.oracle_run.restype = Result
        L.oracle_run.argtypes = [ctypes.c_int64, MatrixData, MatrixData, MatrixData, MethodParams, SpaiExtParams,
                                 ctypes.c_int64]
        L.oracle_free.argtypes = [ctypes.POINTER(Result)]
        L.oracle_product.restype = ctypes.c_int64
        L.oracle_product.argtypes = [MatrixData, MatrixData, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ctypes.c_double)]
        L.oracle_solve_ls2x2.argtypes = [ctypes.POINTER(ctypes.c_double)] * 3
        L.oracle_solve_ls2x2_rank_eps.argtypes = [ctypes.POINTER(ctypes.c_double)] * 3 + [ctypes.c_double]
        L.oracle_run_rank_eps.restype = Result
        L.oracle_run_rank_eps.argtypes = [ctypes.c_int64, MatrixData, MatrixData, MatrixData, MethodParams,
                                          SpaiExtParams, ctypes.c_int64, ctypes.c_double]
        L.oracle_column_topk.restype = ctypes.c_int64
        L.oracle_column_topk.argtypes = [MatrixData, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double)]
        L.oracle_pcg.restype = ctypes.c_int64
        L.oracle_pcg.argtypes = [CsrData, CsrData, ctypes.POINTER(ctypes.c_double), ctypes.c_int64, ctypes.c_double,
                                 ctypes.c_int64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64),
                                 ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def as_csc(matrix) -> sparse.csc_matrix:
    """Canonical CSC with int64 indices, as methods_cython/methods.pyx:74-87 (_as_csc)."""
    csc = sparse.csc_matrix(matrix, dtype=np.float64, copy=True)
    csc.sum_duplicates()
    csc.sort_indices()
    csc.indptr = np.ascontiguousarray(csc.indptr, dtype=np.int64)
    csc.indices = np.ascontiguousarray(csc.indices, dtype=np.int64)
    csc.data = np.ascontiguousarray(csc.data, dtype=np.float64)
    return csc


def _md(csc, keep):
    if csc is None:
        return MatrixData(0, 0, 0, None, None, None)
    keep.append(csc)
    return MatrixData(
        csc.shape[0], csc.shape[1], csc.nnz,
        csc.indptr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
        csc.indices.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
        csc.data.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
    )


def result_to_python(res: Result, free):
    try:
        if res.error_message:
            raise RuntimeError(res.error_message.decode("utf-8", "replace"))
        indptr = np.ctypeslib.as_array(res.col_offsets, shape=(res.cols + 1,)).copy()
        if res.nnz:
            indices = np.ctypeslib.as_array(res.row_indices, shape=(res.nnz,)).copy()
            values = np.ctypeslib.as_array(res.values, shape=(res.nnz,)).copy()
        else:
            indices = np.empty(0, np.int64)
            values = np.empty(0, np.float64)
        M = sparse.csc_matrix((values, indices, indptr), shape=(res.rows, res.cols))
        hist = np.empty(res.history_size, dtype=HISTORY_DTYPE)
        for i in range(res.history_size):
            h = res.history[i]
            hist[i] = (h.iteration, h.residual_norm, h.density_m, h.density_direction)
        return {"M": M, "iterations": int(res.iterations), "converged": bool(res.converged), "history": hist}
    finally:
        free(ctypes.byref(res))


def run(method: str, A, M0=None, Pr=None, max_iterations=100, tolerance=1e-9, max_density=0.03,
        enable_dropping=True, num_threads=1, inner_iterations=2, fixed_pattern=0, column_top_k=0,
        omp_threads=0, rank_eps=None):
    """Runs one method on the CPU oracle; returns dict(M, iterations, converged, history).
    rank_eps: the rank threshold of the inner-outer LOMR 2x2 solves (default
    DBL_EPSILON, the reference's); sqrt(FLT_EPSILON) restates the device's FP32
    form."""
    keep = []
    A_csc = as_csc(A)
    mA = _md(A_csc, keep)
    mM = _md(as_csc(M0), keep) if M0 is not None else _md(None, keep)
    mP = _md(as_csc(Pr), keep) if Pr is not None else _md(None, keep)
    params = MethodParams(int(max_iterations), float(tolerance), float(max_density), 1 if enable_dropping else 0,
                          int(num_threads), 1 if M0 is not None else 0, 1 if Pr is not None else 0,
                          int(inner_iterations))
    ext = SpaiExtParams(int(fixed_pattern), 0, int(column_top_k))
    L = lib()
    if rank_eps is None:
        res = L.oracle_run(METHODS[method], mA, mM, mP, params, ext, int(omp_threads))
    else:
        res = L.oracle_run_rank_eps(METHODS[method], mA, mM, mP, params, ext, int(omp_threads), float(rank_eps))
    return result_to_python(res, L.oracle_free)


def product(X, Y, residual=False, prune_zero=False) -> sparse.csc_matrix:
    """Conservative sparse product X*Y (or I - X*Y with residual=True)."""
    keep = []
    x, y = _md(as_csc(X), keep), _md(as_csc(Y), keep)
    L = lib()
    nnz = L.oracle_product(x, y, int(residual), int(prune_zero), None, None, None)
    if nnz < 0:
        raise RuntimeError("oracle_product failed")
    ptr = np.zeros(Y.shape[1] + 1, np.int64)
    idx = np.zeros(max(nnz, 1), np.int64)
    val = np.zeros(max(nnz, 1), np.float64)
    L.oracle_product(x, y, int(residual), int(prune_zero), ptr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                     idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                     val.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return sparse.csc_matrix((val[:nnz], idx[:nnz], ptr), shape=(X.shape[0], Y.shape[1]))


def column_topk(X, k) -> sparse.csc_matrix:
    """Per-column top-k (diagonal kept) as the oracle's extension applies it."""
    keep = []
    x = _md(as_csc(X), keep)
    L = lib()
    nnz = L.oracle_column_topk(x, int(k), None, None, None)
    ptr = np.zeros(X.shape[1] + 1, np.int64)
    idx = np.zeros(max(nnz, 1), np.int64)
    val = np.zeros(max(nnz, 1), np.float64)
    L.oracle_column_topk(x, int(k), ptr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                         idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                         val.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return sparse.csc_matrix((val[:nnz], idx[:nnz], ptr), shape=X.shape)


def solve_ls2x2(m, x, rank_eps=None):
    m = np.ascontiguousarray(m, dtype=np.float64).reshape(4)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(2)
    out = np.zeros(2)
    dp = ctypes.POINTER(ctypes.c_double)
    if rank_eps is None:
        lib().oracle_solve_ls2x2(m.ctypes.data_as(dp), x.ctypes.data_as(dp), out.ctypes.data_as(dp))
    else:
        lib().oracle_solve_ls2x2_rank_eps(m.ctypes.data_as(dp), x.ctypes.data_as(dp), out.ctypes.data_as(dp),
                                          float(rank_eps))
    return out


def _csr_data(csr, keep):
    csr = sparse.csr_matrix(csr, dtype=np.float64)
    csr.sum_duplicates()
    csr.sort_indices()
    ip = np.ascontiguousarray(csr.indptr, dtype=np.int64)
    ix = np.ascontiguousarray(csr.indices, dtype=np.int64)
    dv = np.ascontiguousarray(csr.data, dtype=np.float64)
    keep.extend([ip, ix, dv])
    return CsrData(csr.shape[0], csr.shape[1], csr.nnz, ip.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                   ix.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), dv.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))


def pcg(A, M, b, max_iterations, tolerance, threads=1):
    """backward_error.pcg restated: returns (history, status, x)."""
    keep = []
    a, m = _csr_data(A, keep), _csr_data(M, keep)
    b = np.ascontiguousarray(b, dtype=np.float64)
    hist = np.zeros(int(max_iterations), np.float64)
    x = np.zeros(b.shape[0], np.float64)
    cnt = ctypes.c_int64(0)
    dp = ctypes.POINTER(ctypes.c_double)
    st = lib().oracle_pcg(a, m, b.ctypes.data_as(dp), int(max_iterations), float(tolerance), int(threads),
                          hist.ctypes.data_as(dp), ctypes.byref(cnt), x.ctypes.data_as(dp))
    if st < 0:
        raise ValueError("n, max_iterations and threads must be positive")
    return hist[:cnt.value].copy(), int(st), x
