"""The no-allocation per-column top-k (spai_column_topk_noalloc) against the
allocating entry point spai_column_topk: bit-identical offsets, rows and
values, on two calls in a row (the second on a smaller matrix, so the shared
device scratch holds the first call's flags), with exact ties in |v|, empty
columns, and columns on both sides of the register path (<= 1024 entries)."""
import ctypes

import numpy as np
import pytest
from scipy import sparse

from parallellocallyoptimalapproximatesparsematrixinverses_amd import _lib

pytestmark = pytest.mark.gpu


def _tied(rows, cols, per_col, seed):
    rng = np.random.default_rng(seed)
    r = np.concatenate([rng.choice(rows, size=min(rows, k), replace=False) for k in per_col])
    c = np.repeat(np.arange(cols), [min(rows, k) for k in per_col])
    X = sparse.csc_matrix((np.round(rng.standard_normal(r.size), 1), (r, c)), shape=(rows, cols))
    X = (X + sparse.identity(max(rows, cols), format="csc")[:rows, :cols]).tocsc()
    X.sum_duplicates()
    X.sort_indices()
    return X


def _declare(L):
    f = L.spai_column_topk_noalloc
    f.restype = ctypes.c_int64
    f.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_uint64]
    return f


def _noalloc(X, k):
    """Runs the no-allocation entry on device copies of X; returns CSC."""
    import torch
    f = _declare(_lib.load())
    X = sparse.csc_matrix(X)
    dev = torch.device("cuda")
    ptr = torch.tensor(X.indptr.astype(np.int32), device=dev)
    idx = torch.tensor(X.indices.astype(np.int32), device=dev)
    val = torch.tensor(X.data.astype(np.float64), device=dev)
    optr = torch.full((X.shape[1] + 1,), -7, dtype=torch.int32, device=dev)
    oidx = torch.full((max(X.nnz, 1),), -7, dtype=torch.int32, device=dev)
    oval = torch.full((max(X.nnz, 1),), np.nan, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    free_before = torch.cuda.mem_get_info(dev)[0]
    rc = f(X.shape[0], X.shape[1], X.nnz, ptr.data_ptr(), idx.data_ptr(), val.data_ptr(), int(k), optr.data_ptr(),
           oidx.data_ptr(), oval.data_ptr(), stream.cuda_stream)
    assert rc == 0, _lib.last_error()
    stream.synchronize()
    free_after = torch.cuda.mem_get_info(dev)[0]
    p = optr.cpu().numpy().astype(np.int64)
    nnz = int(p[-1])
    out = sparse.csc_matrix((oval[:nnz].cpu().numpy(), oidx[:nnz].cpu().numpy().astype(np.int64), p), shape=X.shape)
    return out, free_before - free_after


def _check(X, k):
    ref = _lib.column_topk(X, k)
    got, grown = _noalloc(X, k)
    np.testing.assert_array_equal(got.indptr, ref.indptr)
    np.testing.assert_array_equal(got.indices, ref.indices)
    np.testing.assert_array_equal(got.data, ref.data)
    return grown


def test_noalloc_topk_matches_allocating_entry_two_calls():
    rng = np.random.default_rng(7)
    per_col = rng.integers(0, 400, size=900)
    per_col[::97] = 0           # empty columns (the diagonal alone, or nothing below the square part)
    per_col[5] = 2500           # a column past the register path
    big = _tied(3000, 900, per_col, seed=1)
    small = _tied(700, 500, rng.integers(0, 60, size=500), seed=2)
    print(f"case: noalloc topk big nnz={big.nnz} k=5", flush=True)
    _check(big, 5)
    print(f"case: noalloc topk small nnz={small.nnz} k=3", flush=True)
    grown = _check(small, 3)
    # the scratch array was placed on the device by the first call; a later
    # call takes no device memory
    assert grown == 0, grown


def test_noalloc_topk_keeps_whole_columns_and_empty_matrix():
    X = _tied(50, 40, np.full(40, 3), seed=3)
    print("case: noalloc topk k above every column", flush=True)
    _check(X, 100)
    print("case: noalloc topk k = 0", flush=True)
    _check(X, 0)
    E = sparse.csc_matrix((20, 0))
    got, _ = _noalloc(E, 2)
    assert got.shape == (20, 0) and got.nnz == 0
