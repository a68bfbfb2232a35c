"""Chebyshev connection coefficients and domain-transfer matrices (host, FP64).

Restates ``src/pyloki/utils/maths.py:399-557``:

- ``S[k, m]``: ``x^k = sum_m S[k, m] T_m(x)``, ``2^(1 - k - [m == 0]) binom(k, (k - m)/2)``;
- ``R[k, m]``: ``T_k(x) = sum_m R[k, m] x^m``;
- ``C[n, k](p, q)``: ``T_n(p x + q) = sum_k C[n, k] T_k(x)``, the matrix that moves the
  coefficients of a series on the domain (tc1, ts1) to the domain (tc2, ts2) with
  ``p = ts2 / ts1`` and ``q = (tc2 - tc1) / ts1``.

Float-to-integer powers use binary powering (``psr_utils.powi``), the evaluation order of
every expression follows the reference, and the GPU library computes the same tables in
the same order, so host and device agree bit for bit.
"""

from __future__ import annotations

import math

import numpy as np

from pyloki_amd.utils.psr_utils import powi


def nbinom(n: int, k: int) -> int:
    return math.comb(n, k) if 0 <= k <= n else 0


def compute_connection_coefficient_s(k: int, m: int) -> float:
    if k < 0 or m < 0 or m > k or (k - m) % 2 != 0:
        return 0.0
    n = (k - m) // 2
    delta = 1 if m == 0 else 0
    return 2.0 ** (1 - k - delta) * nbinom(k, n)


def compute_connection_coefficient_r(k: int, m: int) -> float:
    if k < 0 or m < 0 or m > k or (k - m) % 2 != 0:
        return 0.0
    if m == 0 and k % 2 == 0:
        return float((-1) ** (k // 2))
    n = (k - m) // 2
    r = (k + m) // 2
    return ((-1) ** n) * powi(2.0, m - 1) * (2.0 * k / (k + m)) * nbinom(r, n)


def compute_transformation_coefficient_c(n: int, k: int, p: float, q: float) -> float:
    if k > n:
        return 0.0
    result = 0.0
    for m in range(k, n + 1):
        r_nm = compute_connection_coefficient_r(n, m)
        inner = 0.0
        for i in range(k, m + 1):
            inner += (nbinom(m, i) * powi(p, i) * powi(q, m - i)
                      * compute_connection_coefficient_s(i, k))
        result += r_nm * inner
    return result


def compute_connection_matrix_s(k_max: int) -> np.ndarray:
    if k_max < 0:
        msg = "k_max must be a non-negative integer."
        raise ValueError(msg)
    return np.array([[compute_connection_coefficient_s(k, m) for m in range(k_max + 1)]
                     for k in range(k_max + 1)])


def compute_connection_matrix_r(k_max: int) -> np.ndarray:
    return np.array([[compute_connection_coefficient_r(k, m) for m in range(k_max + 1)]
                     for k in range(k_max + 1)])


def poly_chebyshev_transform_matrix(poly_order: int, tc1: float, ts1: float, tc2: float,
                                    ts2: float, coeff_order: int = 0) -> np.ndarray:
    """``C[i, j]``; ``coeff_order`` 0 stores coefficients ascending, 1 descending."""
    if poly_order < 0:
        msg = "k_max must be a non-negative integer."
        raise ValueError(msg)
    if ts1 <= 0 or ts2 <= 0:
        msg = "ts1 and ts2 must be positive."
        raise ValueError(msg)
    p = ts2 / ts1
    q = (tc2 - tc1) / ts1
    size = poly_order + 1
    c = np.zeros((size, size))
    for i in range(size):
        for j in range(size):
            n, k = (i, j) if coeff_order == 0 else (poly_order - i, poly_order - j)
            c[i, j] = compute_transformation_coefficient_c(n, k, p, q)
    return c


def row_times(v: np.ndarray, M: np.ndarray) -> np.ndarray:
    """``v @ M`` for rows of v, each output summed in ascending row order of M."""
    v = np.asarray(v, dtype=np.float64)
    flat = v.reshape(-1, v.shape[-1])
    out = np.zeros((flat.shape[0], M.shape[1]))
    for j in range(M.shape[1]):
        acc = np.zeros(flat.shape[0])
        for i in range(M.shape[0]):
            acc = acc + flat[:, i] * M[i, j]
        out[:, j] = acc
    return out.reshape(*v.shape[:-1], M.shape[1])
