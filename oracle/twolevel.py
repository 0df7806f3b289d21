"""numpy restatement of the two-level PCG preconditioner (libpgn precond 2).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
preconditioner -- never as the product path.

The reference preconditions its CG with Eigen's DiagonalPreconditioner
(source/PhySim/Solvers/LinearSolver_EigenCG.cpp:68-89); libpgn's precond 2
replaces it by M^-1 r = B^-1 r + P Ac^-1 P^T r with B the 3x3 block
diagonal of the Hessian and P the aggregated rigid-body prolongation:
nodes binned by rest position into a grid of boxes shaped to the bounding
box, each aggregate spanning the translations and the infinitesimal
rotations about its centroid over its free DoFs.  What pins this file:
the rigid-body modes annihilate the Hessian of an unstressed, unconstrained
body (an invariant of the energy, checked in tests/test_oracle_twolevel.py),
the box counts follow from the binning rule in closed form, and the inverse
is numpy.linalg.inv.
"""
import numpy as np

MAX_AGGS = 2048 // 6


def bins_for(X0, target=128):
    """Boxes per axis: about `target` of them, proportional to the extents
    (an axis shorter than 1e-3 of the longest counts as that long), at most
    MAX_AGGS in all (the largest count is lowered first)."""
    X = np.asarray(X0, dtype=np.float64).reshape(-1, 3)
    ext = X.max(0) - X.min(0)
    emax = ext.max()
    if emax <= 0:
        return [1, 1, 1]
    e = np.maximum(ext, 1e-3 * emax)
    s = np.cbrt(min(target, MAX_AGGS) / np.prod(e))
    b = [max(1, int(np.floor(s * ek + 0.5))) for ek in e]
    while b[0] * b[1] * b[2] > MAX_AGGS:
        k = 0 if (b[0] >= b[1] and b[0] >= b[2]) else (1 if b[1] >= b[2] else 2)
        b[k] -= 1
    return b


def aggregate(X0, target=128):
    """Aggregate id of every node: its box, boxes numbered x fastest, empty
    boxes dropped."""
    X = np.asarray(X0, dtype=np.float64).reshape(-1, 3)
    b = bins_for(X, target)
    lo, hi = X.min(0), X.max(0)
    ext = hi - lo
    key = np.zeros(len(X), dtype=np.int64)
    idx = []
    for k in range(3):
        if ext[k] > 0:
            v = np.floor((X[:, k] - lo[k]) / ext[k] * b[k]).astype(np.int64)
        else:
            v = np.zeros(len(X), dtype=np.int64)
        idx.append(np.clip(v, 0, b[k] - 1))
    key = (idx[2] * b[1] + idx[1]) * b[0] + idx[0]
    _, agg = np.unique(key, return_inverse=True)
    return agg.astype(np.int32)


def offsets(X0, agg):
    """rel_i = X0_i - centroid of its aggregate."""
    X = np.asarray(X0, dtype=np.float64).reshape(-1, 3)
    na = agg.max() + 1
    cnt = np.bincount(agg, minlength=na)
    cen = np.stack([np.bincount(agg, weights=X[:, k], minlength=na) / cnt for k in range(3)], 1)
    return X - cen[agg]


def node_blocks(rel):
    """P_i (3 x 6) of every node: [I | W(rel_i)], W w = w x rel."""
    n = len(rel)
    P = np.zeros((n, 3, 6))
    P[:, 0, 0] = P[:, 1, 1] = P[:, 2, 2] = 1.0
    x, y, z = rel[:, 0], rel[:, 1], rel[:, 2]
    P[:, 0, 4], P[:, 0, 5] = z, -y
    P[:, 1, 3], P[:, 1, 5] = -z, x
    P[:, 2, 3], P[:, 2, 4] = y, -x
    return P


def kept_columns(X0, agg, fixed):
    """Bit mask of the kept coarse columns per aggregate: incremental
    Cholesky of the Gram matrix of its columns (rotations scaled by the
    aggregate's rms offset), keeping a column whose pivot exceeds 1e-8 of its
    diagonal."""
    rel = offsets(X0, agg)
    na = agg.max() + 1
    cnt = np.bincount(agg, minlength=na)
    rms = np.sqrt(np.bincount(agg, weights=(rel ** 2).sum(1), minlength=na) / cnt)
    sc = np.where(rms > 0, 1.0 / np.where(rms > 0, rms, 1.0), 0.0)
    P = node_blocks(rel * sc[agg][:, None])
    free = 1.0 - np.asarray(fixed, dtype=np.float64).reshape(-1, 3)
    G = np.zeros((na, 6, 6))
    np.add.at(G, agg, np.einsum('npu,np,npv->nuv', P, free, P))
    masks = np.zeros(na, dtype=np.uint8)
    for a in range(na):
        kept, L = [], np.zeros((6, 6))
        for c in range(6):
            gcc = G[a, c, c]
            if not gcc > 0:
                continue
            y = np.zeros(len(kept))
            d = gcc
            for k, kc in enumerate(kept):
                y[k] = (G[a, kc, c] - L[k, :k] @ y[:k]) / L[k, k]
                d -= y[k] * y[k]
            if d > 1e-8 * gcc:
                L[len(kept), :len(kept)] = y
                L[len(kept), len(kept)] = np.sqrt(d)
                kept.append(c)
                masks[a] |= 1 << c
    return masks


def prolongation(X0, agg, fixed, cmask):
    """Dense P (3n x 6 na): rows of fixed DoFs and columns not kept are zero."""
    rel = offsets(X0, agg)
    n = len(rel)
    na = agg.max() + 1
    Pn = node_blocks(rel) * (1.0 - np.asarray(fixed, dtype=np.float64).reshape(-1, 3))[:, :, None]
    P = np.zeros((3 * n, 6 * na))
    for i in range(n):
        P[3 * i:3 * i + 3, 6 * agg[i]:6 * agg[i] + 6] = Pn[i]
    keep = np.array([(cmask[a] >> c) & 1 for a in range(na) for c in range(6)], dtype=bool)
    P[:, ~keep] = 0.0
    return P, keep


def coarse_matrix(A, P, keep):
    """Ac = P^T A P with unit rows / columns for the dropped columns."""
    Ac = P.T @ A @ P
    Ac[~keep, :] = 0.0
    Ac[:, ~keep] = 0.0
    Ac[~keep, ~keep] = 1.0
    return Ac


def block_jacobi(A, r):
    n = len(r) // 3
    z = np.zeros_like(r)
    for i in range(n):
        s = slice(3 * i, 3 * i + 3)
        z[s] = np.linalg.solve(A[s, s], r[s])
    return z


def apply(A, r, P, Ac):
    """M^-1 r = B^-1 r + P Ac^-1 P^T r."""
    return block_jacobi(A, r) + P @ np.linalg.solve(Ac, P.T @ r)
