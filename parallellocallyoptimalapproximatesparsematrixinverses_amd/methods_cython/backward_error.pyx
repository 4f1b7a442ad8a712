# cython: language_level=3
"""PCG evaluator of built preconditioners, backed by libspai_hip.so.

Mirrors the reference module methods_cython.backward_error
(methods_cython/backward_error.pyx:45-144): ``matvec``, ``pcg`` and
``parallel_threads`` keep their signatures; ``threads`` is accepted for
compatibility and the work runs on the current HIP device.  ``pcg`` returns
(relative-residual history, status) with the reference's status codes.
"""
import numpy as np

from libc.stdint cimport int64_t, uintptr_t

from ._capi cimport CsrData, spai_matvec, spai_pcg, spai_last_error


cdef object _i64(object a):
    return np.ascontiguousarray(a, dtype=np.int64)


cdef object _f64(object a):
    return np.ascontiguousarray(a, dtype=np.float64)


cdef CsrData _csr(object values, object indices, object offsets, int64_t n):
    cdef CsrData d
    d.rows = n
    d.cols = n
    d.nnz = <int64_t>values.shape[0]
    d.row_offsets = <const int64_t*><uintptr_t>offsets.ctypes.data
    d.col_indices = <const int64_t*><uintptr_t>indices.ctypes.data
    d.values = <const double*><uintptr_t>values.ctypes.data
    return d


def parallel_threads(Py_ssize_t n, int requested):
    if n <= 0 or requested <= 0:
        raise ValueError('n and requested must be positive')
    return int(min(n, requested))


def matvec(values, indices, offsets, x, int threads):
    values, indices, offsets, x = _f64(values), _i64(indices), _i64(offsets), _f64(x)
    if threads <= 0 or offsets.shape[0] != x.shape[0] + 1:
        raise ValueError('Invalid matrix-vector dimensions or thread count')
    n = x.shape[0]
    out = np.empty(n, dtype=np.float64)
    cdef CsrData a = _csr(values, indices, offsets, n)
    cdef const double* xp = <const double*><uintptr_t>x.ctypes.data
    cdef double* yp = <double*><uintptr_t>out.ctypes.data
    cdef int64_t rc
    with nogil:
        rc = spai_matvec(a, xp, yp)
    if rc < 0:
        raise RuntimeError(spai_last_error().decode("utf-8", "replace"))
    return out


def pcg(a_values, a_indices, a_offsets, m_values, m_indices, m_offsets, b, Py_ssize_t max_iterations, int threads,
        double tolerance, *, m_fp32=False):
    b = _f64(b)
    cdef int64_t n = b.shape[0]
    if n <= 0 or max_iterations <= 0 or threads <= 0:
        raise ValueError('n, max_iterations and threads must be positive')
    a_values, a_indices, a_offsets = _f64(a_values), _i64(a_indices), _i64(a_offsets)
    m_values, m_indices, m_offsets = _f64(m_values), _i64(m_indices), _i64(m_offsets)
    if a_offsets.shape[0] != n + 1 or m_offsets.shape[0] != n + 1:
        raise ValueError('Invalid matrix-vector dimensions')
    cdef CsrData a = _csr(a_values, a_indices, a_offsets, n)
    cdef CsrData m = _csr(m_values, m_indices, m_offsets, n)
    history = np.empty(max_iterations, dtype=np.float64)
    cdef double* hp = <double*><uintptr_t>history.ctypes.data
    cdef const double* bp = <const double*><uintptr_t>b.ctypes.data
    cdef int64_t count = 0
    cdef int64_t status
    cdef int64_t fp32 = 1 if m_fp32 else 0
    with nogil:
        status = spai_pcg(a, m, bp, max_iterations, tolerance, hp, &count, NULL, fp32)
    if status < 0:
        raise RuntimeError(spai_last_error().decode("utf-8", "replace"))
    return history[:count].copy(), int(status)
