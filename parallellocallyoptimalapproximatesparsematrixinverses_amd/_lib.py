"""ctypes bindings of libspai_hip.so (include/methods_c_api.h, include/spai_ext.h).

The library is loaded from this package directory only; a missing library is
an error (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np
from scipy import sparse

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libspai_hip.so"

METHOD_IDS = {"cg": 0, "mr": 1, "lomr": 2, "io_mr": 3, "io_lomr": 4}

C_ENTRY_POINTS = ("spai_global_cg", "spai_global_mr", "spai_global_lomr", "inner_outer_mr", "inner_outer_lomr",
                  "free_mem")
EXT_ENTRY_POINTS = ("spai_run_ex", "spai_pcg", "spai_matvec", "spai_last_error", "spai_device_name",
                    "spai_session_create", "spai_session_destroy", "spai_session_build", "spai_session_set_rhs",
                    "spai_session_pcg", "spai_session_pcg_output", "spai_session_result",
                    "spai_session_pcg_iteration_us")

PCG_LAYOUTS = ("csr", "sell", "dia")

I64P = ctypes.POINTER(ctypes.c_int64)
F64P = ctypes.POINTER(ctypes.c_double)


class MatrixData(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_int64), ("cols", ctypes.c_int64), ("nnz", ctypes.c_int64),
                ("col_offsets", I64P), ("row_indices", I64P), ("values", F64P)]


class MethodParams(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int64), ("tolerance", ctypes.c_double), ("max_density", ctypes.c_double),
                ("enable_dropping", ctypes.c_int64), ("num_threads", ctypes.c_int64), ("use_initial_m", ctypes.c_int64),
                ("use_preconditioner", ctypes.c_int64), ("inner_iterations", ctypes.c_int64)]


class SingleIterationData(ctypes.Structure):
    _fields_ = [("iteration", ctypes.c_int64), ("residual_norm", ctypes.c_double), ("density_m", ctypes.c_double),
                ("density_direction", ctypes.c_double)]


class Result(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_int64), ("cols", ctypes.c_int64), ("nnz", ctypes.c_int64),
                ("iterations", ctypes.c_int64), ("converged", ctypes.c_int64), ("history_size", ctypes.c_int64),
                ("col_offsets", I64P), ("row_indices", I64P), ("values", F64P),
                ("history", ctypes.POINTER(SingleIterationData)), ("error_message", ctypes.c_char_p)]


class SpaiExtParams(ctypes.Structure):
    _fields_ = [("fixed_pattern", ctypes.c_int64), ("precision", ctypes.c_int64), ("column_top_k", ctypes.c_int64),
                ("batch_ordered_reductions", ctypes.c_int64), ("reserved", ctypes.c_int64 * 4)]


class CsrData(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_int64), ("cols", ctypes.c_int64), ("nnz", ctypes.c_int64),
                ("row_offsets", I64P), ("col_indices", I64P), ("values", F64P)]


HISTORY_DTYPE = np.dtype([("iteration", np.int64), ("residual_norm", np.float64), ("density_m", np.float64),
                          ("density_direction", np.float64)])

_lib = None


def load():
    """Loads libspai_hip.so and declares the prototypes (no device call)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(make -C parallellocallyoptimalapproximatesparsematrixinverses_amd/csrc)")
    L = ctypes.CDLL(str(LIB_PATH))
    for name in C_ENTRY_POINTS[:5]:
        f = getattr(L, name)
        f.restype = Result
        f.argtypes = [MatrixData, MatrixData, MatrixData, MethodParams]
    L.free_mem.restype = None
    L.free_mem.argtypes = [ctypes.POINTER(Result)]
    L.spai_run_ex.restype = Result
    L.spai_run_ex.argtypes = [ctypes.c_int64, MatrixData, MatrixData, MatrixData, MethodParams, SpaiExtParams, F64P]
    L.spai_spgemm.restype = Result
    L.spai_spgemm.argtypes = [MatrixData, MatrixData, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, F64P]
    L.spai_column_topk.restype = Result
    L.spai_column_topk.argtypes = [MatrixData, ctypes.c_int64]
    L.spai_pcg.restype = ctypes.c_int64
    L.spai_pcg.argtypes = [CsrData, CsrData, F64P, ctypes.c_int64, ctypes.c_double, F64P, I64P, F64P, ctypes.c_int64]
    L.spai_matvec.restype = ctypes.c_int64
    L.spai_matvec.argtypes = [CsrData, F64P, F64P]
    L.spai_last_error.restype = ctypes.c_char_p
    L.spai_last_error.argtypes = []
    L.spai_device_name.restype = ctypes.c_int64
    L.spai_device_name.argtypes = [ctypes.c_char_p, ctypes.c_int64]
    L.spai_session_create.restype = ctypes.c_void_p
    L.spai_session_create.argtypes = [ctypes.c_int64, MatrixData, MatrixData, MatrixData, MethodParams, SpaiExtParams]
    L.spai_session_destroy.restype = None
    L.spai_session_destroy.argtypes = [ctypes.c_void_p]
    L.spai_session_build.restype = ctypes.c_int64
    L.spai_session_build.argtypes = [ctypes.c_void_p]
    L.spai_session_set_rhs.restype = ctypes.c_int64
    L.spai_session_set_rhs.argtypes = [ctypes.c_void_p, F64P]
    L.spai_session_pcg.restype = ctypes.c_int64
    L.spai_session_pcg.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, I64P]
    L.spai_session_pcg_output.restype = ctypes.c_int64
    L.spai_session_pcg_output.argtypes = [ctypes.c_void_p, F64P, F64P]
    L.spai_session_result.restype = Result
    L.spai_session_result.argtypes = [ctypes.c_void_p]
    L.spai_session_pcg_iteration_us.restype = ctypes.c_double
    L.spai_session_pcg_iteration_us.argtypes = [ctypes.c_void_p]
    L.spai_session_pcg_layouts.restype = ctypes.c_int64
    L.spai_session_pcg_layouts.argtypes = [ctypes.c_void_p, I64P]
    L.spai_session_set_kernel_timing.restype = ctypes.c_int64
    L.spai_session_set_kernel_timing.argtypes = [ctypes.c_void_p, ctypes.c_int64]
    L.spai_session_build_kernel_us.restype = ctypes.c_double
    L.spai_session_build_kernel_us.argtypes = [ctypes.c_void_p, I64P]
    L.spai_stream_bandwidth.restype = ctypes.c_int64
    L.spai_stream_bandwidth.argtypes = [ctypes.c_int64, ctypes.c_int64, F64P]
    L.spai_session_info.restype = ctypes.c_int64
    L.spai_session_info.argtypes = [ctypes.c_void_p, I64P, ctypes.c_int64]
    _lib = L
    return L


def last_error() -> str:
    return load().spai_last_error().decode("utf-8", "replace")


def stream_bandwidth(bytes_per_array=2 << 30, repeats=20) -> dict:
    """Achievable HBM GB/s of the device: streaming read, copy and triad."""
    out = (ctypes.c_double * 3)()
    if load().spai_stream_bandwidth(int(bytes_per_array), int(repeats), out) < 0:
        raise RuntimeError(last_error())
    return {"read": out[0], "copy": out[1], "triad": out[2]}


def device_name() -> str:
    L = load()
    buf = ctypes.create_string_buffer(256)
    if L.spai_device_name(buf, 256) < 0:
        raise RuntimeError(last_error())
    return buf.value.decode()


def _is_canonical(m) -> bool:
    return (sparse.isspmatrix_csc(m) and m.dtype == np.float64 and m.indptr.dtype == np.int64
            and m.indices.dtype == np.int64 and m.indptr.flags.c_contiguous and m.indices.flags.c_contiguous
            and m.data.flags.c_contiguous and m.has_canonical_format)


def as_csc(matrix, shape=None) -> sparse.csc_matrix:
    """Canonical CSC with int64 indices (methods_cython/methods.pyx:74-87
    semantics).  A matrix already in that form is used as it is (the library
    only reads it); anything else is copied and canonicalised."""
    if _is_canonical(matrix):
        if shape is not None and matrix.shape != tuple(shape):
            raise ValueError("matrix shape does not match A")
        return matrix
    csc = sparse.csc_matrix(matrix, dtype=np.float64, copy=True)
    if shape is not None and csc.shape != tuple(shape):
        raise ValueError("matrix shape does not match A")
    csc.sum_duplicates()
    csc.sort_indices()
    csc.indptr = np.ascontiguousarray(csc.indptr, dtype=np.int64)
    csc.indices = np.ascontiguousarray(csc.indices, dtype=np.int64)
    csc.data = np.ascontiguousarray(csc.data, dtype=np.float64)
    return csc


def matrix_data(csc) -> MatrixData:
    if csc is None:
        return MatrixData(0, 0, 0, None, None, None)
    return MatrixData(csc.shape[0], csc.shape[1], csc.nnz, csc.indptr.ctypes.data_as(I64P),
                      csc.indices.ctypes.data_as(I64P), csc.data.ctypes.data_as(F64P))


def csr_data(csr, keep: list) -> CsrData:
    csr = sparse.csr_matrix(csr, dtype=np.float64)
    csr.sum_duplicates()
    csr.sort_indices()
    ip = np.ascontiguousarray(csr.indptr, dtype=np.int64)
    ix = np.ascontiguousarray(csr.indices, dtype=np.int64)
    dv = np.ascontiguousarray(csr.data, dtype=np.float64)
    keep.extend([ip, ix, dv])
    return CsrData(csr.shape[0], csr.shape[1], csr.nnz, ip.ctypes.data_as(I64P), ix.ctypes.data_as(I64P),
                   dv.ctypes.data_as(F64P))


def result_to_python(res: Result) -> dict:
    """Copies a Result into numpy/scipy objects and releases it with free_mem."""
    L = load()
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
        L.free_mem(ctypes.byref(res))


def _params(max_iterations, tolerance, max_density, enable_dropping, num_threads, use_m0, use_pr, inner_iterations):
    return MethodParams(int(max_iterations), float(tolerance), float(max_density), 1 if enable_dropping else 0,
                        int(num_threads), 1 if use_m0 else 0, 1 if use_pr else 0, int(inner_iterations))


def run(method: str, A, M0=None, Pr=None, max_iterations=100, tolerance=1e-9, max_density=0.03,
        enable_dropping=True, num_threads=1, inner_iterations=2, fixed_pattern=False, precision="fp64",
        column_top_k=0, batch_ordered=False, plain_entry=False):
    """Runs a builder on the device.  plain_entry=True calls the reference-named
    entry point (spai_global_* / inner_outer_*) instead of spai_run_ex."""
    L = load()
    A_csc = as_csc(A)
    M0_csc = as_csc(M0, A_csc.shape) if M0 is not None else None
    Pr_csc = as_csc(Pr, A_csc.shape) if Pr is not None else None
    params = _params(max_iterations, tolerance, max_density, enable_dropping, num_threads, M0 is not None,
                     Pr is not None, inner_iterations)
    args = (matrix_data(A_csc), matrix_data(M0_csc), matrix_data(Pr_csc), params)
    seconds = ctypes.c_double(0.0)
    if plain_entry:
        name = {"cg": "spai_global_cg", "mr": "spai_global_mr", "lomr": "spai_global_lomr",
                "io_mr": "inner_outer_mr", "io_lomr": "inner_outer_lomr"}[method]
        res = getattr(L, name)(*args)
    else:
        ext = SpaiExtParams(1 if fixed_pattern else 0, 1 if precision == "fp32" else 0, int(column_top_k),
                            1 if batch_ordered else 0)
        res = L.spai_run_ex(METHOD_IDS[method], *args, ext, ctypes.byref(seconds))
    out = result_to_python(res)
    out["device_seconds"] = seconds.value
    return out


def spgemm(X, Y, residual=False, prune_zero=False, repeats=1):
    """C = X Y (or I - X Y) by the device SpGEMM of the reference-semantics
    path; returns (C as CSC, mean seconds of one product over `repeats`)."""
    L = load()
    X_csc, Y_csc = as_csc(X), as_csc(Y)
    seconds = ctypes.c_double(0.0)
    res = L.spai_spgemm(matrix_data(X_csc), matrix_data(Y_csc), 1 if residual else 0, 1 if prune_zero else 0,
                        int(repeats), ctypes.byref(seconds))
    out = result_to_python(res)
    return out["M"], seconds.value


def column_topk(X, k):
    """Per-column top-k by |v| (ties: smaller row first), diagonal kept, on
    the device; returns CSC."""
    L = load()
    X_csc = as_csc(X)  # keeps the arrays matrix_data points into alive across the call
    res = L.spai_column_topk(matrix_data(X_csc), int(k))
    return result_to_python(res)["M"]


def pcg(A, M, b, max_iterations, tolerance, m_fp32=False):
    """Device PCG; returns (history, status, x)."""
    L = load()
    keep = []
    a, m = csr_data(A, keep), csr_data(M, keep)
    b = np.ascontiguousarray(b, dtype=np.float64)
    hist = np.zeros(int(max_iterations), np.float64)
    x = np.zeros(b.shape[0], np.float64)
    cnt = ctypes.c_int64(0)
    st = L.spai_pcg(a, m, b.ctypes.data_as(F64P), int(max_iterations), float(tolerance), hist.ctypes.data_as(F64P),
                    ctypes.byref(cnt), x.ctypes.data_as(F64P), 1 if m_fp32 else 0)
    if st < 0:
        raise ValueError(last_error())
    return hist[:cnt.value].copy(), int(st), x


def matvec(A, x):
    L = load()
    keep = []
    a = csr_data(A, keep)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(A.shape[0], np.float64)
    if L.spai_matvec(a, x.ctypes.data_as(F64P), y.ctypes.data_as(F64P)) < 0:
        raise ValueError(last_error())
    return y


class Session:
    """Device-resident build + PCG for one A (fixed-pattern forms)."""

    def __init__(self, method, A, M0=None, Pr=None, max_iterations=30, tolerance=0.0, max_density=1.0,
                 num_threads=1, precision="fp64", inner_iterations=2):
        L = load()
        self._keep = [as_csc(A), as_csc(M0) if M0 is not None else None, as_csc(Pr) if Pr is not None else None]
        self.n = A.shape[0]
        params = _params(max_iterations, tolerance, max_density, False, num_threads, M0 is not None, Pr is not None,
                         inner_iterations)
        ext = SpaiExtParams(1, 1 if precision == "fp32" else 0, 0)
        self._h = L.spai_session_create(METHOD_IDS[method], matrix_data(self._keep[0]), matrix_data(self._keep[1]),
                                        matrix_data(self._keep[2]), params, ext)
        if not self._h:
            raise RuntimeError(last_error())

    def build(self) -> int:
        it = load().spai_session_build(self._h)
        if it < 0:
            raise RuntimeError(last_error())
        return int(it)

    def set_rhs(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        if load().spai_session_set_rhs(self._h, b.ctypes.data_as(F64P)) < 0:
            raise RuntimeError(last_error())

    def pcg(self, max_iterations, tolerance):
        cnt = ctypes.c_int64(0)
        st = load().spai_session_pcg(self._h, int(max_iterations), float(tolerance), ctypes.byref(cnt))
        if st < 0:
            raise RuntimeError(last_error())
        return int(st), int(cnt.value)

    def pcg_output(self, count):
        hist = np.zeros(count, np.float64)
        x = np.zeros(self.n, np.float64)
        if load().spai_session_pcg_output(self._h, hist.ctypes.data_as(F64P), x.ctypes.data_as(F64P)) < 0:
            raise RuntimeError(last_error())
        return hist, x

    def result(self) -> dict:
        return result_to_python(load().spai_session_result(self._h))

    def download_result(self) -> int:
        """The C-ABI result call alone (M and the history copied to host
        arrays, then released), without building Python objects; returns
        nnz(M)."""
        L = load()
        res = L.spai_session_result(self._h)
        try:
            if res.error_message:
                raise RuntimeError(res.error_message.decode("utf-8", "replace"))
            return int(res.nnz)
        finally:
            L.free_mem(ctypes.byref(res))

    def spmv_a_us(self) -> float:
        v = load().spai_session_pcg_iteration_us(self._h)
        if v < 0:
            raise RuntimeError(last_error())
        return float(v)

    def pcg_layouts(self):
        """Row layouts of the PCG iteration kernels for (A, M): "csr", "sell" or "dia"."""
        out = (ctypes.c_int64 * 2)()
        if load().spai_session_pcg_layouts(self._h, out) < 0:
            raise RuntimeError(last_error())
        return tuple(PCG_LAYOUTS[int(v)] for v in out)

    def set_kernel_timing(self, on=True):
        if load().spai_session_set_kernel_timing(self._h, 1 if on else 0) < 0:
            raise RuntimeError(last_error())

    def build_kernel_us(self):
        """(average microseconds, passes) of the timed per-iteration pass of the last build."""
        n = ctypes.c_int64(0)
        us = load().spai_session_build_kernel_us(self._h, ctypes.byref(n))
        return float(us), int(n.value)

    def info(self) -> dict:
        """Launch shape of the build: per-iteration pass (1 fused, 2
        diagonal, 0 neither), lanes per column, columns per tile, tiles,
        resident-grid blocks, wavefronts per block; nnz of the patterns S
        (M's) and S2 (pattern(A S) | S | I); diagonals of A, S and S2 on the
        diagonal pass (0 otherwise)."""
        keys = ("fused", "lanes_per_column", "columns_per_tile", "tiles", "grid_blocks", "waves_per_block", "nnz_s",
                "nnz_s2", "diagonals_a", "diagonals_s", "diagonals_s2")
        out = (ctypes.c_int64 * len(keys))()
        if load().spai_session_info(self._h, out, len(keys)) < 0:
            raise RuntimeError(last_error())
        return {k: int(v) for k, v in zip(keys, out)}

    def close(self):
        if self._h:
            load().spai_session_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
