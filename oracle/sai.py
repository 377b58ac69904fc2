"""Sparse approximate inverse smoother, restated in numpy (FP64).

Test infrastructure only (see oracle/__init__.py).  No reference counterpart:
the reference offers Jacobi and mass scaling (core/bem_solver.py:36-117); this
restates the native left SAI so its rows can be checked:

  row i of M minimises |m^T A[J, I] - e_i^T|, m supported on J(i);
  J(i) = [i] + the (32 - 1) nearest panels j with |c_i - c_j| <= f_J (R_i + R_j),
  I(i) = [i] + the (128 - 1) nearest with f_I (ties by tree index);
  A_ab = K_E A_b / sqrt(max(r_ab^2, 1e-24)) + delta_ab (near correction),
  A_aa = self_corr_a A_a;
  normal equations (B^T B + lam I) m = B^T e_i, B[l][k] = A[J_k][I_l],
  lam = 1e-12 trace(B^T B) / |J|.

Parity: pinned in tests by the exact-inverse case (J = I = every panel of a
20-panel mesh gives M = A^-1) and by |I - M A|_F falling below Jacobi's.
"""
import math

import numpy as np

from .constants import K_E
from .quadrature import near_pairs

CAP_J, CAP_I, CAND = 32, 128, 512


def operator_entries(C, A, self_corr, near=None):
    """Dense A (n x n) of the collocation operator with optional near
    correction ``near`` = (pairs (P, 2), weights (P,))."""
    C = np.asarray(C, dtype=np.float64)
    A = np.asarray(A, dtype=np.float64)
    D = C[:, None, :] - C[None, :, :]
    r2 = np.maximum((D ** 2).sum(-1), 1e-24)
    M = K_E * A[None, :] / np.sqrt(r2)
    np.fill_diagonal(M, np.asarray(self_corr) * A)
    if near is not None and len(near[0]):
        p, w = near
        np.add.at(M, (p[:, 0], p[:, 1]), w)
    return M


def _select(i, cands, C, cap, tree_rank):
    cands = list(cands)[:CAND]
    if len(cands) <= cap - 1:
        return [i] + cands
    d = [float(((C[i] - C[j]) ** 2).sum()) for j in cands]
    order = sorted(range(len(cands)), key=lambda t: (d[t], tree_rank[cands[t]]))
    return [i] + [cands[t] for t in order[:cap - 1]]


def sai_matrix(C, A, self_corr, near=None, factor_pattern=2.0, factor_rows=4.0, perm=None):
    """Dense M (n x n).  ``perm`` (tree order -> original index) sets the tie
    order of the nearest-candidate selection and the candidate order."""
    C = np.asarray(C, dtype=np.float64)
    n = C.shape[0]
    perm = np.arange(n) if perm is None else np.asarray(perm)
    rank = np.empty(n, dtype=np.int64)
    rank[perm] = np.arange(n)
    Aop = operator_entries(C, A, self_corr, near)
    pj = near_pairs(C, A, factor_pattern)
    pi = near_pairs(C, A, factor_rows)
    rowsJ = [[] for _ in range(n)]
    rowsI = [[] for _ in range(n)]
    for a, b in pj:
        rowsJ[a].append(b)
    for a, b in pi:
        rowsI[a].append(b)
    M = np.zeros((n, n))
    for i in range(n):
        J = _select(i, sorted(rowsJ[i], key=lambda j: rank[j]), C, CAP_J, rank)
        I = _select(i, sorted(rowsI[i], key=lambda j: rank[j]), C, CAP_I, rank)
        B = Aop[np.ix_(J, I)].T
        G = B.T @ B
        lam = 1e-12 * (np.trace(G) if np.trace(G) > 0 else 1.0) / len(J)
        m = np.linalg.solve(G + lam * np.eye(len(J)), B[0])
        M[i, J] = m
    return M


_ = math
