"""Argument errors of the extension entry points, reported before any device
work (so they run without a GPU): the reference's contract of a zeroed
Result with error_message, surfaced by the Python bindings as RuntimeError."""
import numpy as np
import pytest
from scipy import sparse

from parallellocallyoptimalapproximatesparsematrixinverses_amd import _lib


def _m(rows, cols):
    return sparse.random(rows, cols, density=0.2, random_state=0, format="csc")


def test_spgemm_argument_errors():
    with pytest.raises(RuntimeError, match="dimensions do not match"):
        _lib.spgemm(_m(4, 3), _m(4, 4))
    with pytest.raises(RuntimeError, match="square result"):
        _lib.spgemm(_m(5, 4), _m(4, 3), residual=True)
    with pytest.raises(RuntimeError, match="repeats must be positive"):
        _lib.spgemm(_m(4, 4), _m(4, 4), repeats=0)


def test_column_topk_argument_errors():
    with pytest.raises(RuntimeError, match="non-negative"):
        _lib.column_topk(_m(5, 5), -1)


def test_malformed_csc_is_rejected():
    X = _m(6, 6)
    bad = _CscView(X.shape, X.indptr[::-1], X.indices, X.data)  # decreasing column offsets
    L = _lib.load()
    res = L.spai_column_topk(_lib.matrix_data(bad), 2)
    try:
        assert res.error_message and b"non-decreasing" in res.error_message
        assert res.nnz == 0 and res.rows == 0
    finally:
        L.free_mem(res)


class _CscView:
    """CSC arrays as int64/float64 without scipy's checks (to pass malformed
    offsets through to the library)."""

    def __init__(self, shape, indptr, indices, data):
        self.shape = shape
        self.nnz = len(data)
        self.indptr = np.ascontiguousarray(indptr, dtype=np.int64)
        self.indices = np.ascontiguousarray(indices, dtype=np.int64)
        self.data = np.ascontiguousarray(data, dtype=np.float64)


def test_session_entry_points_reject_null_session():
    """Every session entry point answers a null session with -1 and a message
    (no host crash), before any device work."""
    import ctypes
    L = _lib.load()
    f64 = ctypes.POINTER(ctypes.c_double)
    buf = (ctypes.c_double * 4)()
    cnt = ctypes.c_int64(0)
    out = (ctypes.c_int64 * 6)()
    calls = [
        lambda: L.spai_session_build(None),
        lambda: L.spai_session_set_rhs(None, ctypes.cast(buf, f64)),
        lambda: L.spai_session_pcg(None, 10, 1e-8, ctypes.byref(cnt)),
        lambda: L.spai_session_pcg_output(None, ctypes.cast(buf, f64), ctypes.cast(buf, f64)),
        lambda: L.spai_session_info(None, out, 6),
    ]
    for call in calls:
        assert call() == -1
        assert _lib.last_error() in ("null session", "session has no built preconditioner")
    assert L.spai_session_pcg_iteration_us(None) < 0
    assert _lib.last_error() == "session has no built preconditioner"
    assert L.spai_session_build_kernel_us(None, ctypes.byref(cnt)) < 0


def test_column_topk_noalloc_rejects_sizes_beyond_its_scratch():
    """The no-allocation top-k refuses inputs beyond its fixed device scratch
    (16,777,216 entries, 4,194,304 columns) before touching the device."""
    import ctypes
    L = _lib.load()
    f = L.spai_column_topk_noalloc
    f.restype = ctypes.c_int64
    f.argtypes = [ctypes.c_int64] * 3 + [ctypes.c_void_p] * 3 + [ctypes.c_int64] + [ctypes.c_void_p] * 3 + [
        ctypes.c_uint64]
    dummy = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    assert f(10, 10, (1 << 24) + 1, dummy, dummy, dummy, 2, dummy, dummy, dummy, 0) == -1
    assert "exceeds the maximum 16777216" in _lib.last_error()
    assert f(10, (1 << 22) + 1, 10, dummy, dummy, dummy, 2, dummy, dummy, dummy, 0) == -1
    assert "exceeds the maximum 4194304" in _lib.last_error()
    assert f(10, 10, 10, dummy, dummy, dummy, -1, dummy, dummy, dummy, 0) == -1
    assert "k must be non-negative" in _lib.last_error()
