"""Laplace FMM expansions and translation operators (numpy, FP64).

Test infrastructure only (see oracle/__init__.py).

This is the convention of the native FMM, not a restatement of a reference
file: the reference builds its translation matrices per pair by least squares
(fmm3d/multipole_operators.py:665-981), which is not the parity target (the
dense matvec core/bem_kernel.py:157-230 is).  The functions here are pinned
directly against the mathematics in tests: 1/|x-y| reproduced by the
expansion, P2M->M2L->L2P and P2M->M2M / L2L->L2P against direct summation with
the truncation bound, and the real-packed matrices against the complex formulas.

Solid harmonics (Condon-Shortley associated Legendre P_n^m):
    R_n^m(x) = r^n P_n^m(cos t) e^{i m phi} / (n+m)!
    S_n^m(x) = (n-m)! P_n^m(cos t) e^{i m phi} / r^(n+1)
    X_n^{-m} = (-1)^m conj(X_n^m)   for X in {R, S}
    1/|x-y| = sum_{n,m} conj(R_n^m(y)) S_n^m(x),  |y| < |x|
Expansions about a centre c with box scale s (all translation maps then depend
only on offsets in units of s):
    P2M  M_n^m = sum_i q_i conj(R_n^m((y_i - c)/s))
    M2P  phi(x) = (1/s) sum M_n^m S_n^m((x - c)/s)
    M2M  M'_n^m = sum_{j,k} M_j^k 2^-j conj(R_{n-j}^{m-k}((c - c')/s'))   (child s = s'/2)
    M2L  L_n^m = (1/s) (-1)^(n+m) sum_{j,k} M_j^k S_{j+n}^{k-m}((c_L - c_M)/s)
    L2L  L'_j^k = sum_{n,m} L_n^m 2^-j R_{n-j}^{m-k}((c' - c)/s)            (child s' = s/2)
    L2P  phi(x) = sum L_n^m R_n^m((x - c)/s)
Real packing of the m >= 0 half: index n^2 for Re X_n^0, n^2+2m-1 / n^2+2m for
Re / Im X_n^m (m >= 1); (p+1)^2 reals in all.
"""
import math

import numpy as np


def n_coef(p):
    return (p + 1) * (p + 1)


def regular(x, p):
    """R[n, m] (0<=m<=n<=p) at points x (..., 3); complex array (..., p+1, p+1)."""
    x = np.asarray(x, dtype=float)
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    r2 = X * X + Y * Y + Z * Z
    R = np.zeros(x.shape[:-1] + (p + 1, p + 1), dtype=complex)
    R[..., 0, 0] = 1.0
    w = X + 1j * Y
    for m in range(1, p + 1):
        R[..., m, m] = -w / (2 * m) * R[..., m - 1, m - 1]
    for m in range(0, p):
        R[..., m + 1, m] = Z * R[..., m, m]
        for n in range(m + 1, p):
            R[..., n + 1, m] = ((2 * n + 1) * Z * R[..., n, m] - r2 * R[..., n - 1, m]) / (
                (n + 1) ** 2 - m * m)
    return R


def irregular(x, p):
    """S[n, m] (0<=m<=n<=p) at points x (..., 3)."""
    x = np.asarray(x, dtype=float)
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    r2 = X * X + Y * Y + Z * Z
    inv_r2 = 1.0 / r2
    S = np.zeros(x.shape[:-1] + (p + 1, p + 1), dtype=complex)
    S[..., 0, 0] = 1.0 / np.sqrt(r2)
    w = X + 1j * Y
    for m in range(1, p + 1):
        S[..., m, m] = -(2 * m - 1) * w * inv_r2 * S[..., m - 1, m - 1]
    for m in range(0, p):
        S[..., m + 1, m] = (2 * m + 1) * Z * inv_r2 * S[..., m, m]
        for n in range(m + 1, p):
            S[..., n + 1, m] = ((2 * n + 1) * Z * S[..., n, m]
                                - (n * n - m * m) * S[..., n - 1, m]) * inv_r2
    return S


def full(Xh, n, m):
    """X_n^m for any m in [-n, n] from the m >= 0 table (last two axes)."""
    if abs(m) > n:
        return 0.0
    if m >= 0:
        return Xh[..., n, m]
    return (-1) ** (-m) * np.conj(Xh[..., n, -m])


def pack(Xh, p):
    out = np.zeros(Xh.shape[:-2] + (n_coef(p),))
    for n in range(p + 1):
        out[..., n * n] = Xh[..., n, 0].real
        for m in range(1, n + 1):
            out[..., n * n + 2 * m - 1] = Xh[..., n, m].real
            out[..., n * n + 2 * m] = Xh[..., n, m].imag
    return out


def unpack(v, p):
    v = np.asarray(v, dtype=float)
    Xh = np.zeros(v.shape[:-1] + (p + 1, p + 1), dtype=complex)
    for n in range(p + 1):
        Xh[..., n, 0] = v[..., n * n]
        for m in range(1, n + 1):
            Xh[..., n, m] = v[..., n * n + 2 * m - 1] + 1j * v[..., n * n + 2 * m]
    return Xh


def p2m(points, q, center, s, p):
    R = regular((np.asarray(points) - center) / s, p)
    return pack(np.einsum("i,inm->nm", np.asarray(q, dtype=float), np.conj(R)), p)


def m2p(Mv, targets, center, s, p):
    M = unpack(Mv, p)
    S = irregular((np.asarray(targets) - center) / s, p)
    out = np.zeros(S.shape[0])
    for n in range(p + 1):
        out += (M[n, 0] * S[:, n, 0]).real
        for m in range(1, n + 1):
            out += 2.0 * (M[n, m] * S[:, n, m]).real
    return out / s


def l2p(Lv, targets, center, s, p):
    L = unpack(Lv, p)
    R = regular((np.asarray(targets) - center) / s, p)
    out = np.zeros(R.shape[0])
    for n in range(p + 1):
        out += (L[n, 0] * R[:, n, 0]).real
        for m in range(1, n + 1):
            out += 2.0 * (L[n, m] * R[:, n, m]).real
    return out


def _apply_complex(fn, p_in, p_out):
    """Real matrix of a real-linear map given on complex m>=0 tables."""
    nin, nout = n_coef(p_in), n_coef(p_out)
    T = np.zeros((nout, nin))
    for col in range(nin):
        e = np.zeros(nin)
        e[col] = 1.0
        T[:, col] = pack(fn(unpack(e, p_in)), p_out)
    return T


def m2l_matrix(offset, p):
    """Real (p+1)^2 x (p+1)^2 M2L matrix for an offset (c_L - c_M)/s; the
    caller multiplies by 1/s."""
    S = irregular(np.asarray(offset, dtype=float), 2 * p)

    def fn(M):
        L = np.zeros((p + 1, p + 1), dtype=complex)
        for n in range(p + 1):
            for m in range(0, n + 1):
                acc = 0.0
                for j in range(p + 1):
                    for k in range(-j, j + 1):
                        acc += full(M, j, k) * full(S, j + n, k - m)
                L[n, m] = (-1) ** (n + m) * acc
        return L

    return _apply_complex(fn, p, p)


def m2m_matrix(d, p):
    """Child->parent M2M for d = (c_child - c_parent)/s_parent, child scale
    s_parent/2."""
    R = regular(np.asarray(d, dtype=float), p)

    def fn(M):
        out = np.zeros((p + 1, p + 1), dtype=complex)
        for n in range(p + 1):
            for m in range(0, n + 1):
                acc = 0.0
                for j in range(n + 1):
                    for k in range(-j, j + 1):
                        acc += full(M, j, k) * 2.0 ** (-j) * np.conj(full(R, n - j, m - k))
                out[n, m] = acc
        return out

    return _apply_complex(fn, p, p)


def l2l_matrix(d, p):
    """Parent->child L2L for d = (c_child - c_parent)/s_parent, child scale
    s_parent/2."""
    R = regular(np.asarray(d, dtype=float), p)

    def fn(L):
        out = np.zeros((p + 1, p + 1), dtype=complex)
        for j in range(p + 1):
            for k in range(0, j + 1):
                acc = 0.0
                for n in range(j, p + 1):
                    for m in range(-n, n + 1):
                        acc += full(L, n, m) * 2.0 ** (-j) * full(R, n - j, m - k)
                out[j, k] = acc
        return out

    return _apply_complex(fn, p, p)


OCTANT_SIGNS = np.array([[(o >> 0) & 1, (o >> 1) & 1, (o >> 2) & 1] for o in range(8)],
                        dtype=float) * 2.0 - 1.0


def child_offset(octant):
    """(c_child - c_parent)/s_parent for octant bits x | y<<1 | z<<2."""
    return 0.25 * OCTANT_SIGNS[octant]


def truncation_bound(rho_src, dist, p):
    """Relative M2L truncation scale ((a+b)/d)^(p+1)."""
    return (rho_src / dist) ** (p + 1)


def factorial(n):
    return math.factorial(n)
