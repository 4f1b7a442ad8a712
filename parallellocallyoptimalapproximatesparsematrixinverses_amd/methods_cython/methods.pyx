# cython: language_level=3
"""Python API of the SPAI builders, backed by libspai_hip.so.

Mirrors the reference module methods_cython.methods (methods_cython/methods.pyx:182-258):
``global_spai.{cg,mr,lomr}`` and ``inner_outer.{mr,lomr}`` take scipy matrices
and return ``OutputResult(M, iterations, converged, history)``.  Each call
runs on the current HIP device with the GIL released.  Keyword-only extras
select the forms the reference lacks: ``fixed_pattern``, ``precision`` ("fp64" or
"fp32") and ``column_top_k``.
"""
import numpy as np
from scipy import sparse

from libc.stdint cimport int64_t, uintptr_t
from libc.string cimport memset

from ._capi cimport (MatrixData, MethodParams, Result, SpaiExtParams, free_mem, spai_global_cg, spai_global_mr,
                     spai_global_lomr, inner_outer_mr, inner_outer_lomr, spai_run_ex)

ctypedef Result (*entry_t)(MatrixData, MatrixData, MatrixData, MethodParams) noexcept nogil

HISTORY_DTYPE = np.dtype([
    ("iteration", np.int64),
    ("residual_norm", np.float64),
    ("density_m", np.float64),
    ("density_direction", np.float64),
])

# method id -> (reference-named entry point, spai_run_ex id)
cdef int _CG = 0, _MR = 1, _LOMR = 2, _IO_MR = 3, _IO_LOMR = 4


class OutputResult:
    """Built preconditioner M (scipy CSC) with its per-iteration history."""

    def __init__(self, M, iterations, converged, history):
        self.M = M
        self.iterations = iterations
        self.converged = converged
        self.history = history
        self.residual_norms = history["residual_norm"]


def _canonical_csc(matrix, shape=None):
    if matrix is None:
        return sparse.csc_matrix((0, 0), dtype=np.float64)
    out = sparse.csc_matrix(matrix, dtype=np.float64, copy=True)
    if shape is not None and out.shape != tuple(shape):
        raise ValueError("matrix shape does not match A")
    out.sum_duplicates()
    out.sort_indices()
    out.indptr = np.ascontiguousarray(out.indptr, dtype=np.int64)
    out.indices = np.ascontiguousarray(out.indices, dtype=np.int64)
    out.data = np.ascontiguousarray(out.data, dtype=np.float64)
    return out


cdef MatrixData _borrow(object csc):
    cdef MatrixData d
    d.rows = <int64_t>csc.shape[0]
    d.cols = <int64_t>csc.shape[1]
    d.nnz = <int64_t>csc.nnz
    d.col_offsets = <const int64_t*><uintptr_t>csc.indptr.ctypes.data
    d.row_indices = <const int64_t*><uintptr_t>csc.indices.ctypes.data
    d.values = <const double*><uintptr_t>csc.data.ctypes.data
    return d


cdef object _take(Result res):
    cdef int64_t i
    try:
        if res.error_message != NULL:
            raise RuntimeError((<bytes>res.error_message).decode("utf-8", "replace"))
        indptr = np.asarray(<int64_t[:res.cols + 1]>res.col_offsets).copy()
        if res.nnz > 0:
            indices = np.asarray(<int64_t[:res.nnz]>res.row_indices).copy()
            values = np.asarray(<double[:res.nnz]>res.values).copy()
        else:
            indices = np.empty(0, dtype=np.int64)
            values = np.empty(0, dtype=np.float64)
        history = np.empty(res.history_size, dtype=HISTORY_DTYPE)
        for i in range(res.history_size):
            history[i] = (res.history[i].iteration, res.history[i].residual_norm,
                          res.history[i].density_m, res.history[i].density_direction)
        M = sparse.csc_matrix((values, indices, indptr), shape=(res.rows, res.cols))
        return OutputResult(M, res.iterations, bool(res.converged), history)
    finally:
        free_mem(&res)


cdef object _run(int method, object A, object M0, object Pr, int64_t max_iterations, double tolerance,
                 double max_density, object enable_dropping, int64_t num_threads, int64_t inner_iterations,
                 object fixed_pattern, object precision, int64_t column_top_k):
    if precision not in ("fp64", "fp32"):
        raise ValueError("precision must be 'fp64' or 'fp32'")
    a = _canonical_csc(A)
    m0 = _canonical_csc(M0, a.shape) if M0 is not None else _canonical_csc(None)
    pr = _canonical_csc(Pr, a.shape) if Pr is not None else _canonical_csc(None)
    cdef MatrixData da = _borrow(a)
    cdef MatrixData dm = _borrow(m0)
    cdef MatrixData dp = _borrow(pr)
    cdef MethodParams params
    params.max_iterations = max_iterations
    params.tolerance = tolerance
    params.max_density = max_density
    params.enable_dropping = 1 if enable_dropping else 0
    params.num_threads = num_threads
    params.use_initial_m = 0 if M0 is None else 1
    params.use_preconditioner = 0 if Pr is None else 1
    params.inner_iterations = inner_iterations
    cdef SpaiExtParams ext
    memset(&ext, 0, sizeof(ext))
    ext.fixed_pattern = 1 if fixed_pattern else 0
    ext.precision = 1 if precision == "fp32" else 0
    ext.column_top_k = column_top_k
    cdef bint extended = ext.fixed_pattern != 0 or ext.precision != 0 or ext.column_top_k != 0
    cdef entry_t entry
    if method == _CG:
        entry = spai_global_cg
    elif method == _MR:
        entry = spai_global_mr
    elif method == _LOMR:
        entry = spai_global_lomr
    elif method == _IO_MR:
        entry = inner_outer_mr
    else:
        entry = inner_outer_lomr
    cdef Result res
    with nogil:
        if extended:
            res = spai_run_ex(method, da, dm, dp, params, ext, NULL)
        else:
            res = entry(da, dm, dp, params)
    return _take(res)


class GlobalSpaiMethods:
    available_methods = ("cg", "mr", "lomr")

    def cg(self, A, M0=None, Pr=None, max_iterations=100, tolerance=1e-9, max_density=0.03, enable_dropping=True,
           num_threads=1, *, fixed_pattern=False, precision="fp64", column_top_k=0):
        return _run(_CG, A, M0, Pr, max_iterations, tolerance, max_density, enable_dropping, num_threads, 2,
                    fixed_pattern, precision, column_top_k)

    def mr(self, A, M0=None, Pr=None, max_iterations=100, tolerance=1e-9, max_density=0.03, enable_dropping=True,
           num_threads=1, *, fixed_pattern=False, precision="fp64", column_top_k=0):
        return _run(_MR, A, M0, Pr, max_iterations, tolerance, max_density, enable_dropping, num_threads, 2,
                    fixed_pattern, precision, column_top_k)

    def lomr(self, A, M0=None, Pr=None, max_iterations=100, tolerance=1e-9, max_density=0.03, enable_dropping=True,
             num_threads=1, *, fixed_pattern=False, precision="fp64", column_top_k=0):
        return _run(_LOMR, A, M0, Pr, max_iterations, tolerance, max_density, enable_dropping, num_threads, 2,
                    fixed_pattern, precision, column_top_k)


class InnerOuterMethods:
    available_methods = ("mr", "lomr")

    def mr(self, A, M0=None, Pr=None, max_iterations=100, tolerance=1e-9, max_density=0.03, enable_dropping=True,
           num_threads=1, inner_iterations=2, *, fixed_pattern=False, precision="fp64", column_top_k=0):
        return _run(_IO_MR, A, M0, Pr, max_iterations, tolerance, max_density, enable_dropping, num_threads,
                    inner_iterations, fixed_pattern, precision, column_top_k)

    def lomr(self, A, M0=None, Pr=None, max_iterations=100, tolerance=1e-9, max_density=0.03, enable_dropping=True,
             num_threads=1, inner_iterations=2, *, fixed_pattern=False, precision="fp64", column_top_k=0):
        return _run(_IO_LOMR, A, M0, Pr, max_iterations, tolerance, max_density, enable_dropping, num_threads,
                    inner_iterations, fixed_pattern, precision, column_top_k)


global_spai = GlobalSpaiMethods()
inner_outer = InnerOuterMethods()
