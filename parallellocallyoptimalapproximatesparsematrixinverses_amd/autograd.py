"""Differentiation form of the SPAI objective on fixed patterns.

    f(M, A) = ||I - A M||_F^2,   R = I - A M
    df/dM = -2 (A^T R) sampled on pattern(M)
    df/dA = -2 (R M^T) sampled on pattern(A)

``ResidualLoss(A_pattern, M_pattern, dtype, block_cols)`` holds the device
plan (patterns, transposes, blocking); calling it with the value vectors of M
and A (torch tensors on the HIP device, in the CSC order of the patterns)
returns the loss as a 0-d FP64 tensor through ``FrobeniusResidual``, a
``torch.autograd.Function`` whose backward runs the library's HIP kernels.
FP32 values use FP32 products with FP64 reductions and gradient accumulation.
R is formed per block of ``block_cols`` columns and never stored whole.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib


def _load():
    L = _lib.load()
    if not hasattr(L, "_loss_declared"):
        L.spai_loss_plan_create.restype = ctypes.c_void_p
        L.spai_loss_plan_create.argtypes = [_lib.MatrixData, _lib.MatrixData, ctypes.c_int64, ctypes.c_int64,
                                            ctypes.c_uint64]
        L.spai_loss_plan_destroy.restype = None
        L.spai_loss_plan_destroy.argtypes = [ctypes.c_void_p]
        L.spai_loss_set_stream.restype = ctypes.c_int64
        L.spai_loss_set_stream.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        L.spai_loss_plan_info.restype = ctypes.c_int64
        L.spai_loss_plan_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
        L.spai_loss_forward.restype = ctypes.c_int64
        L.spai_loss_forward.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.POINTER(ctypes.c_double)]
        L.spai_loss_backward.restype = ctypes.c_int64
        L.spai_loss_backward.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                         ctypes.c_void_p, ctypes.c_void_p]
        L.spai_loss_last_error.restype = ctypes.c_char_p
        L.spai_loss_last_error.argtypes = []
        L._loss_declared = True
    return L


def _error(L):
    return L.spai_loss_last_error().decode("utf-8", "replace")


class ResidualLoss:
    """Device plan of f(M, A) = ||I - A M||_F^2 for fixed patterns of A and M."""

    def __init__(self, A_pattern, M_pattern, dtype=torch.float64, block_cols=0, device="cuda"):
        if dtype not in (torch.float32, torch.float64):
            raise ValueError("dtype must be torch.float32 or torch.float64")
        L = _load()
        self.dtype = dtype
        device = torch.device(device)
        if device.type != "cuda":
            raise ValueError("the plan runs on a HIP device")
        # the plan's device memory lives on the device current at creation
        self.device = torch.device("cuda", device.index if device.index is not None else torch.cuda.current_device())
        self.A = _lib.as_csc(A_pattern)
        self.M = _lib.as_csc(M_pattern, self.A.shape)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        self._h = L.spai_loss_plan_create(_lib.matrix_data(self.A), _lib.matrix_data(self.M),
                                          1 if dtype == torch.float32 else 0, int(block_cols), int(stream))
        if not self._h:
            raise RuntimeError(_error(L))

    def info(self):
        out = (ctypes.c_int64 * 4)()
        _load().spai_loss_plan_info(self._h, out)
        return {"nnz_r": out[0], "max_block_entries": out[1], "blocks": out[2], "nnz_m": out[3]}

    def _check(self, t, n):
        if t.dtype != self.dtype or t.device != self.device or t.numel() != n or not t.is_contiguous():
            raise ValueError(f"expected a contiguous {self.dtype} tensor of {n} values on {self.device}")

    def _use_current_stream(self, L):
        # the kernels run on the caller's current stream, ordered after the
        # work that produced the value tensors
        if L.spai_loss_set_stream(self._h, torch.cuda.current_stream(self.device).cuda_stream) < 0:
            raise RuntimeError(_error(L))

    def forward_value(self, m_values, a_values) -> float:
        self._check(m_values, self.M.nnz)
        self._check(a_values, self.A.nnz)
        L = _load()
        self._use_current_stream(L)
        out = ctypes.c_double(0.0)
        if L.spai_loss_forward(self._h, m_values.data_ptr(), a_values.data_ptr(), ctypes.byref(out)) < 0:
            raise RuntimeError(_error(L))
        return out.value

    def gradients(self, m_values, a_values, scale=1.0):
        self._check(m_values, self.M.nnz)
        self._check(a_values, self.A.nnz)
        L = _load()
        self._use_current_stream(L)
        gm = torch.empty_like(m_values)
        ga = torch.empty_like(a_values)
        if L.spai_loss_backward(self._h, m_values.data_ptr(), a_values.data_ptr(), float(scale), gm.data_ptr(),
                                ga.data_ptr()) < 0:
            raise RuntimeError(_error(L))
        return gm, ga

    def __call__(self, m_values, a_values):
        return FrobeniusResidual.apply(m_values, a_values, self)

    def close(self):
        if getattr(self, "_h", None):
            _load().spai_loss_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FrobeniusResidual(torch.autograd.Function):
    """||I - A M||_F^2 with hand-written HIP backward kernels."""

    @staticmethod
    def forward(ctx, m_values, a_values, plan: ResidualLoss):
        ctx.plan = plan
        ctx.save_for_backward(m_values, a_values)
        value = plan.forward_value(m_values.detach(), a_values.detach())
        return torch.tensor(value, dtype=torch.float64, device=m_values.device)

    @staticmethod
    def backward(ctx, grad_out):
        m_values, a_values = ctx.saved_tensors
        scale = float(grad_out.detach().cpu())
        gm, ga = ctx.plan.gradients(m_values.detach(), a_values.detach(), scale)
        return (gm if ctx.needs_input_grad[0] else None, ga if ctx.needs_input_grad[1] else None, None)
