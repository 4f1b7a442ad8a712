"""Synthetic SPD test problems (all matrices CSC, sorted, symmetric).

  poisson2d(32)              2D 5-point Laplacian, Dirichlet (4 / -1); n=1,024, nnz=4,992
  diffusion3d(64, seed=1)    3D 7-point variable-coefficient diffusion on 64^3: cell
                             coefficients k = exp(g), g ~ N(0,1) (numpy default_rng(seed)),
                             face coefficients 2 k_i k_j / (k_i + k_j); a face on the domain
                             boundary carries the cell's own k (Dirichlet ghost of equal
                             coefficient).  n=262,144, nnz=1,810,432
  random_gram(100000, seed=3) A = B^T B + sigma I, B with 6 N(0,1) entries per row at uniform
                             columns, sigma = 1e-4 * lambda_max(B^T B) from 50 power iterations
  anisotropic2d(1024)        2D 5-point, coefficient 1 in x and 1e-3 in y, Dirichlet
  stencil27(128, seed=4)     3D 27-point, off-diagonals -exp(g) (symmetric, one g per pair),
                             diagonal = sum |off-diagonal| + 1e-3
"""
from __future__ import annotations

import numpy as np
from scipy import sparse


def _canonical(A) -> sparse.csc_matrix:
    A = sparse.csc_matrix(A, dtype=np.float64)
    A.sum_duplicates()
    A.sort_indices()
    A.indptr = A.indptr.astype(np.int64)
    A.indices = A.indices.astype(np.int64)
    return A


def poisson2d(m: int = 32) -> sparse.csc_matrix:
    T = sparse.diags([-np.ones(m - 1), 2 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])
    I = sparse.identity(m)
    return _canonical(sparse.kron(I, T) + sparse.kron(T, I))


def _faces(shape):
    """Index pairs (i, j) of neighbouring cells along each axis (C order)."""
    idx = np.arange(int(np.prod(shape))).reshape(shape)
    pairs = []
    for ax in range(len(shape)):
        a = np.take(idx, np.arange(shape[ax] - 1), axis=ax).ravel()
        b = np.take(idx, np.arange(1, shape[ax]), axis=ax).ravel()
        pairs.append((a, b))
    return pairs


def diffusion3d(m: int = 64, seed: int = 1) -> sparse.csc_matrix:
    shape = (m, m, m)
    n = m ** 3
    k = np.exp(np.random.default_rng(seed).standard_normal(n))
    rows, cols, vals = [], [], []
    diag = np.zeros(n)
    interior_faces = np.zeros(n)
    for a, b in _faces(shape):
        c = 2.0 * k[a] * k[b] / (k[a] + k[b])
        rows += [a, b]
        cols += [b, a]
        vals += [-c, -c]
        np.add.at(diag, a, c)
        np.add.at(diag, b, c)
        np.add.at(interior_faces, a, 1)
        np.add.at(interior_faces, b, 1)
    diag += (6 - interior_faces) * k  # boundary faces
    rows.append(np.arange(n))
    cols.append(np.arange(n))
    vals.append(diag)
    A = sparse.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    return _canonical(A)


def anisotropic2d(m: int = 1024, cx: float = 1.0, cy: float = 1e-3) -> sparse.csc_matrix:
    T = sparse.diags([-np.ones(m - 1), 2 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])
    I = sparse.identity(m)
    return _canonical(cx * sparse.kron(I, T) + cy * sparse.kron(T, I))


def random_gram(n: int = 100_000, per_row: int = 6, seed: int = 3, power_iterations: int = 50):
    rng = np.random.default_rng(seed)
    cols = rng.integers(0, n, size=(n, per_row))
    vals = rng.standard_normal((n, per_row))
    B = sparse.csr_matrix((vals.ravel(), cols.ravel(), np.arange(0, n * per_row + 1, per_row)), shape=(n, n))
    B.sum_duplicates()
    G = (B.T @ B).tocsc()
    v = np.ones(n) / np.sqrt(n)
    lam = 0.0
    for _ in range(power_iterations):
        w = G @ v
        lam = float(np.linalg.norm(w))
        v = w / lam
    sigma = 1e-4 * lam
    return _canonical(G + sigma * sparse.identity(n)), sigma, lam


def stencil27(m: int = 128, seed: int = 4) -> sparse.csc_matrix:
    n = m ** 3
    idx = np.arange(n).reshape(m, m, m)
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    # the 13 "forward" neighbour offsets of the 26-neighbourhood
    offsets = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)
               if (dx, dy, dz) > (0, 0, 0)]
    for dx, dy, dz in offsets:
        sl_a = tuple(slice(max(0, -d), m - max(0, d)) for d in (dx, dy, dz))
        sl_b = tuple(slice(max(0, d), m - max(0, -d)) for d in (dx, dy, dz))
        rows.append(idx[sl_a].ravel())
        cols.append(idx[sl_b].ravel())
    a = np.concatenate(rows)
    b = np.concatenate(cols)
    w = -np.exp(rng.standard_normal(a.shape[0]))
    diag = np.zeros(n)
    np.add.at(diag, a, -w)
    np.add.at(diag, b, -w)
    diag += 1e-3
    A = sparse.coo_matrix((np.concatenate([w, w, diag]),
                           (np.concatenate([a, b, np.arange(n)]), np.concatenate([b, a, np.arange(n)]))),
                          shape=(n, n))
    return _canonical(A)


def diag_inverse(A) -> sparse.csc_matrix:
    """Pr = diag(A)^-1 as the reference harness builds it (scripts/common.py:402-409)."""
    d = A.diagonal()
    if not np.isfinite(d).all() or np.any(d <= 0.0):
        raise ValueError("A must have a finite positive diagonal.")
    return _canonical(sparse.diags(1.0 / d))
