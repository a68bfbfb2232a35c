"""Host-side Taylor coordinate transforms used by planning (FP64).

The per-leaf transforms run on the GPU; these host versions serve ``ParamLimits`` and
other planning code, restating ``src/pyloki/utils/transforms.py``:

- ``shift_taylor_params``: ``transforms.py:10-47``. The matrix entry is
  ``dt**p / p!`` with ``dt**p`` by binary powering (``psr_utils.powi``), and the
  product is summed in ascending column order.
- ``taylor_to_cheby``: ``transforms.py:435-480`` (Chebyshev planning bounds).
"""

from __future__ import annotations

import math

import numpy as np

from pyloki_amd.utils.psr_utils import powi


def taylor_shift_matrix(n_params: int, delta_t: float) -> np.ndarray:
    """Lower-triangular matrix ``T[i, j] = dt^(i-j) / (i-j)!`` for ``i >= j``."""
    t_mat = np.zeros((n_params, n_params), dtype=np.float64)
    for i in range(n_params):
        for j in range(i + 1):
            p = i - j
            t_mat[i, j] = powi(delta_t, p) / float(math.factorial(p))
    return t_mat


def shift_taylor_params(taylor_param_vec: np.ndarray, delta_t: float, n_out: int = 0) -> np.ndarray:
    """Shift ``[d_k, ..., d_1, d_0]`` to a reference time ``delta_t`` later."""
    vec = np.asarray(taylor_param_vec, dtype=np.float64)
    n_params = vec.shape[-1]
    n_out = n_params if n_out <= 0 else min(n_out, n_params)
    t_mat = taylor_shift_matrix(n_params, delta_t)[-n_out:]
    flat = vec.reshape(-1, n_params)
    out = np.zeros((flat.shape[0], n_out), dtype=np.float64)
    for r in range(flat.shape[0]):
        for i in range(n_out):
            acc = 0.0
            for j in range(n_params):
                acc += flat[r, j] * t_mat[i, j]
            out[r, i] = acc
    return out.reshape(*vec.shape[:-1], n_out)


def shift_taylor_errors(taylor_error_vec: np.ndarray, delta_t: float,
                        tiling_strategy: str = "aggressive") -> np.ndarray:
    """Propagate cell widths ``[dd_k .. dd_0]`` by dt (``transforms.py:50-90``)."""
    err = np.asarray(taylor_error_vec, dtype=np.float64)
    n = err.shape[-1]
    t_mat = taylor_shift_matrix(n, delta_t)
    flat = err.reshape(-1, n)
    if tiling_strategy == "aggressive":
        out = flat * np.abs(np.diag(t_mat))[None, :]
    elif tiling_strategy == "conservative":
        out = np.zeros_like(flat)
        for i in range(n):
            for j in range(n):
                out[:, i] = out[:, i] + flat[:, j] * abs(t_mat[i, j])
    elif tiling_strategy == "quadrature":
        out = np.zeros_like(flat)
        for i in range(n):
            for j in range(n):
                out[:, i] = out[:, i] + flat[:, j] ** 2 * t_mat[i, j] ** 2
        out = np.sqrt(out)
    else:
        msg = f"Invalid tiling strategy: {tiling_strategy}"
        raise ValueError(msg)
    return out.reshape(err.shape)


def taylor_to_cheby(taylor_vec: np.ndarray, t_s: float) -> np.ndarray:
    """Taylor ``[d_k, ..., d_0]`` to Chebyshev ``[alpha_k, ..., alpha_0]`` on [-t_s, t_s].

    ``transforms.py:435-463``: scale ``d_k t_s^k / k!`` then convert the power basis
    to the Chebyshev basis (the connection matrix S of ``maths.py``).
    """
    if t_s <= 0:
        msg = "t_s must be a positive."
        raise ValueError(msg)
    d = np.asarray(taylor_vec, dtype=np.float64)[::-1]
    n = len(d)
    power = np.array([d[k] * t_s**k / math.factorial(k) for k in range(n)])
    alpha = np.zeros(n, dtype=np.float64)
    conv = np.polynomial.chebyshev.poly2cheb(power)
    alpha[: len(conv)] = conv
    return alpha[::-1]
