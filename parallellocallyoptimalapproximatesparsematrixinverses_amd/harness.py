"""Backward-error harness on the device: the caller-side API of
scripts/utils/backward_error.py (random right-hand sides, prepared PCG
system and solver, status strings), with A and M held in HBM by a prepared
system of the C ABI (spai_pcg_prepare / spai_pcg_solve / spai_pcg_matvec).

Names, defaults, argument meaning and error behaviour follow the reference:
``random_rhs_realizations`` draws ``rng.random(n)`` from ``default_rng(seed)``
and multiplies by A with row dots in column order (so b is bit-identical to
the reference's); ``PreparedPCGSolver.calculate`` uses ``limit =
max_iterations or n`` and tolerance 1e-6, and maps status codes to the same
strings.  ``num_threads`` is accepted for signature compatibility; the device
decides its own parallelism."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
from scipy import sparse

from . import _lib

DEFAULT_MAX_ITERATIONS = 50
DEFAULT_REALIZATIONS = 20
TOLERANCE = 1e-6

STATUS_STRINGS = {
    0: "iteration limit reached",
    1: "converged",
    2: "PCG breakdown (zero or nonfinite recurrence scalar)",
    3: "PCG breakdown (nonfinite residual)",
}


def _declare(L):
    if getattr(L, "_harness_declared", False):
        return L
    L.spai_pcg_prepare.restype = ctypes.c_void_p
    L.spai_pcg_prepare.argtypes = [_lib.CsrData, _lib.CsrData, ctypes.c_int64]
    L.spai_pcg_release.restype = None
    L.spai_pcg_release.argtypes = [ctypes.c_void_p]
    L.spai_pcg_solve.restype = ctypes.c_int64
    L.spai_pcg_solve.argtypes = [ctypes.c_void_p, _lib.F64P, ctypes.c_int64, ctypes.c_double, _lib.F64P,
                                 ctypes.POINTER(ctypes.c_int64), _lib.F64P]
    L.spai_pcg_matvec.restype = ctypes.c_int64
    L.spai_pcg_matvec.argtypes = [ctypes.c_void_p, _lib.F64P, _lib.F64P]
    L.spai_pcg_layouts.restype = ctypes.c_int64
    L.spai_pcg_layouts.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
    L._harness_declared = True
    return L


def _prepare_matrix(matrix) -> sparse.csr_matrix:
    """CSR, FP64, canonical (sorted, no duplicates), finite values
    (scripts/utils/backward_error.py:47-60)."""
    matrix = sparse.csr_matrix(matrix, dtype=np.float64, copy=False)
    if (not matrix.has_canonical_format or not matrix.data.flags.c_contiguous
            or not matrix.indices.flags.c_contiguous or not matrix.indptr.flags.c_contiguous):
        matrix = matrix.copy()
        matrix.sum_duplicates()
    matrix.check_format(full_check=True)
    if not np.isfinite(matrix.data).all():
        raise ValueError("A and M must contain finite values")
    return matrix


@dataclass
class BackwardErrorResult:
    history: np.ndarray
    converged: bool
    status: str
    solver: str = "PCG"


class _DeviceSystem:
    """Owns one prepared system (A and M on the device)."""

    def __init__(self, A: sparse.csr_matrix, M: sparse.csr_matrix, m_fp32: bool = False):
        L = _declare(_lib.load())
        self._keep = []
        self.n = A.shape[0]
        self._h = L.spai_pcg_prepare(_lib.csr_data(A, self._keep), _lib.csr_data(M, self._keep), 1 if m_fp32 else 0)
        if not self._h:
            raise ValueError(_lib.last_error())

    def solve(self, b, limit, tolerance):
        L = _declare(_lib.load())
        b = np.ascontiguousarray(b, dtype=np.float64)
        hist = np.zeros(int(limit), np.float64)
        x = np.zeros(self.n, np.float64)
        cnt = ctypes.c_int64(0)
        st = L.spai_pcg_solve(self._h, b.ctypes.data_as(_lib.F64P), int(limit), float(tolerance),
                              hist.ctypes.data_as(_lib.F64P), ctypes.byref(cnt), x.ctypes.data_as(_lib.F64P))
        if st < 0:
            raise ValueError(_lib.last_error())
        return hist[:cnt.value].copy(), int(st), x

    def matvec(self, x):
        L = _declare(_lib.load())
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n, np.float64)
        if L.spai_pcg_matvec(self._h, x.ctypes.data_as(_lib.F64P), y.ctypes.data_as(_lib.F64P)) < 0:
            raise ValueError(_lib.last_error())
        return y

    def layouts(self):
        """Row layouts of the PCG iteration kernels for (A, M): "csr", "sell" or "dia"."""
        out = (ctypes.c_int64 * 2)()
        if _declare(_lib.load()).spai_pcg_layouts(self._h, out) < 0:
            raise ValueError(_lib.last_error())
        return tuple(_lib.PCG_LAYOUTS[int(v)] for v in out)

    def close(self):
        if getattr(self, "_h", None):
            _declare(_lib.load()).spai_pcg_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def random_rhs(A, seed: int = 1, num_threads: int = 1) -> np.ndarray:
    return next(random_rhs_realizations(A, seed=seed, realizations=1, num_threads=num_threads))


def random_rhs_realizations(A, seed: int = 1, realizations: int = DEFAULT_REALIZATIONS, num_threads: int = 1):
    """b = A u with u = rng.random(n) from default_rng(seed), redrawn while
    ||b|| = 0 (scripts/utils/backward_error.py:21-36)."""
    A = _prepare_matrix(A)
    n = A.shape[0]
    if n <= 0 or A.shape[1] != n or A.nnz == 0:
        raise ValueError("A must be a nonempty, nonzero square matrix")
    if num_threads <= 0:
        raise ValueError("num_threads must be positive")
    if realizations <= 0:
        raise ValueError("realizations must be positive")
    system = _DeviceSystem(A, sparse.identity(n, format="csr"))
    rng = np.random.default_rng(seed)
    try:
        for _ in range(realizations):
            b = system.matvec(rng.random(n))
            while np.linalg.norm(b) == 0:
                b = system.matvec(rng.random(n))
            yield b
    finally:
        system.close()


class PreparedPCGSystem:
    """A held for repeated preconditioned solves (scripts/utils/backward_error.py:63-74)."""

    def __init__(self, A, num_threads: int = 1):
        if A.shape[0] == 0 or A.shape[0] != A.shape[1]:
            raise ValueError("A must be a nonempty square matrix")
        if num_threads <= 0 or num_threads > np.iinfo(np.int32).max:
            raise ValueError("num_threads must be a positive OpenMP thread count")
        self.A = _prepare_matrix(A)
        self.num_threads = min(num_threads, A.shape[0])
        self.active_threads = self.num_threads

    def prepare_preconditioner(self, M, m_fp32: bool = False) -> "PreparedPCGSolver":
        return PreparedPCGSolver(self, M, m_fp32=m_fp32)


class PreparedPCGSolver:
    """A and M on the device; one PCG per right-hand side
    (scripts/utils/backward_error.py:77-96)."""

    def __init__(self, system: PreparedPCGSystem, M, m_fp32: bool = False):
        if M.shape != system.A.shape:
            raise ValueError("A and M must have the same shape")
        self.system = system
        self.M = _prepare_matrix(M)
        self._device = _DeviceSystem(system.A, self.M, m_fp32=m_fp32)
        self.last_x = None

    def calculate(self, max_iterations: int = DEFAULT_MAX_ITERATIONS, *, b=None,
                  tolerance: float = TOLERANCE) -> BackwardErrorResult:
        if max_iterations < 0:
            raise ValueError("max_iterations must be nonnegative")
        A = self.system.A
        b = random_rhs(A, num_threads=self.system.num_threads) if b is None else np.ascontiguousarray(b, np.float64)
        b_norm = np.linalg.norm(b)
        if b.shape != (A.shape[0],) or not np.isfinite(b_norm) or b_norm <= 0:
            raise ValueError("b must be a finite vector with ||b|| > 0")
        limit = max_iterations or A.shape[0]
        history, status, x = self._device.solve(b, limit, tolerance)
        self.last_x = x
        return BackwardErrorResult(history, status == 1, STATUS_STRINGS[status])

    def close(self):
        self._device.close()


def calculate_backward_error(A, M, max_iterations: int = DEFAULT_MAX_ITERATIONS, *, b=None,
                             num_threads: int = 1) -> BackwardErrorResult:
    system = PreparedPCGSystem(A, num_threads=num_threads)
    solver = system.prepare_preconditioner(M)
    try:
        return solver.calculate(max_iterations, b=b)
    finally:
        solver.close()
