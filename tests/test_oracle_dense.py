"""Dense pure-Python restatement of the no-dropping global methods, against the
oracle's sparse implementation (n <= 30).

The dense form keeps no sparsity pattern: products sum over k ascending with
Python floats (the reference's product order), batch partial sums are
accumulated as src/common/methods_common.cpp:148-154 does, and "nnz" is the
count of nonzero entries, which equals the reference's nnz because every
no-dropping operation prunes exact zeros.  Agreement is exact.
"""
import math

import numpy as np
import pytest
from scipy import sparse

from helpers import tri


def dense(A):
    return [list(map(float, row)) for row in np.asarray(A.todense())]


def matmul(X, Y):
    n = len(X)
    out = [[0.0] * n for _ in range(n)]
    for j in range(n):
        for i in range(n):
            s = None
            for k in range(n):
                if X[i][k] != 0.0 and Y[k][j] != 0.0:
                    s = X[i][k] * Y[k][j] if s is None else s + X[i][k] * Y[k][j]
            out[i][j] = 0.0 if s is None else s
    return out


def batches(n, b):
    return [(n * i // b, n * (i + 1) // b) for i in range(b)]


def bdot(X, Y, bs):
    tot = 0.0
    for lo, hi in bs:
        s = 0.0
        for j in range(lo, hi):
            for i in range(len(X)):
                if X[i][j] != 0.0 and Y[i][j] != 0.0:
                    s += X[i][j] * Y[i][j]
        tot += s
    return tot


def residual(A, M):
    AM = matmul(A, M)
    n = len(A)
    return [[(1.0 if i == j else 0.0) - AM[i][j] if i == j else -AM[i][j] for j in range(n)] for i in range(n)]


def axpy(X, a, Y):
    return [[x + a * y for x, y in zip(rx, ry)] for rx, ry in zip(X, Y)]


def nnz(X):
    return sum(1 for r in X for v in r if v != 0.0)


def dense_mr(A, M, Pr, iters, bs):
    # src/global_spai/minimal_residual.cpp:31-93 without dropping
    n = len(A)
    R = residual(A, M)
    hist = [math.sqrt(bdot(R, R, bs))]
    Z = matmul(Pr, R)
    for _ in range(iters):
        AZ = matmul(A, Z)
        T = matmul(Pr, AZ)
        num, den = bdot(Z, T, bs), bdot(T, T, bs)
        alpha = num / den
        M = axpy(M, alpha, Z)
        R = residual(A, M)
        hist.append(math.sqrt(bdot(R, R, bs)))
        Z = matmul(Pr, R)
    return M, hist


def dense_cg(A, M, Pr, iters, bs):
    # src/global_spai/conjugate_gradient.cpp:15-83 without dropping
    R = residual(A, M)
    hist = [math.sqrt(bdot(R, R, bs))]
    Z = matmul(Pr, R)
    P = [row[:] for row in Z]
    r_z = bdot(R, Z, bs)
    AP = matmul(A, P)
    for _ in range(iters):
        den = bdot(P, AP, bs)
        alpha = r_z / den
        prev = r_z
        M = axpy(M, alpha, P)
        R = axpy(R, -alpha, AP)
        hist.append(math.sqrt(bdot(R, R, bs)))
        Z = matmul(Pr, R)
        r_z = bdot(R, Z, bs)
        beta = r_z / prev
        P = [[p * beta + z for p, z in zip(rp, rz)] for rp, rz in zip(P, Z)]
        AP = matmul(A, P)
    return M, hist


@pytest.mark.parametrize("threads", [1, 3])
@pytest.mark.parametrize("method", ["mr", "cg"])
def test_dense_restatement_matches(oracle_lib, method, threads):
    n = 12
    A = tri(n)
    Pr = sparse.diags(np.full(n, 0.5), format="csc")
    I = sparse.identity(n, format="csc")
    iters = 5
    r = oracle_lib.run(method, A, I, Pr, iters, 0.0, 1.0, False, threads)
    f = dense_mr if method == "mr" else dense_cg
    M, hist = f(dense(A), dense(I), dense(Pr), iters, batches(n, threads))
    np.testing.assert_array_equal(r["history"]["residual_norm"], np.array(hist))
    np.testing.assert_array_equal(r["M"].toarray(), np.array(M))
    assert r["M"].nnz == nnz(M)
