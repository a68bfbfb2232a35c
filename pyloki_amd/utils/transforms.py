"""Host-side Taylor coordinate transforms used by planning (FP64).

The per-leaf transforms run on the GPU; these host versions serve ``ParamLimits`` and
other planning code, restating ``src/pyloki/utils/transforms.py``:

- ``shift_taylor_params``: ``transforms.py:10-47``. The matrix entry is
  ``dt**p / p!`` with ``dt**p`` by binary powering (``psr_utils.powi``), and the
  product is summed in ascending column order.
- Chebyshev basis: ``transforms.py:435-737`` (``taylor_to_cheby[_full/_errors]``,
  ``cheby_to_taylor[_full]``, ``shift_cheby_full/_errors``, ``cheby_to_taylor_param_shift``)
  with the tables of ``utils/maths.py``; rows times a matrix are summed in ascending order
  (``maths.row_times``), as the GPU library does for the same products.
"""

from __future__ import annotations

import math

import numpy as np

from pyloki_amd.utils import maths
from pyloki_amd.utils.psr_utils import fact, powi


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


def _check_ts(t_s: float) -> None:
    if t_s <= 0:
        msg = "t_s must be a positive."
        raise ValueError(msg)


def _power_scale(k_max: int, t_s: float) -> np.ndarray:
    return np.array([powi(t_s, k) / fact(k) for k in range(k_max + 1)])


def taylor_to_cheby(taylor_vec: np.ndarray, t_s: float) -> np.ndarray:
    """Taylor ``[d_k, ..., d_0]`` to Chebyshev ``[alpha_k, ..., alpha_0]`` on [-t_s, t_s]
    (``transforms.py:435-463``): ``(d_k t_s^k / k!) @ S``."""
    _check_ts(t_s)
    v = np.asarray(taylor_vec, dtype=np.float64)
    k_max = v.shape[-1] - 1
    s_mat = maths.compute_connection_matrix_s(k_max)
    return maths.row_times(v[..., ::-1] * _power_scale(k_max, t_s), s_mat)[..., ::-1]


def taylor_to_cheby_full(taylor_full_vec: np.ndarray, t_s: float) -> np.ndarray:
    """Values and cell widths ``[..., k+1, 2]`` (``transforms.py:466-500``); widths add
    in quadrature through S."""
    _check_ts(t_s)
    v = np.asarray(taylor_full_vec, dtype=np.float64)
    k_max = v.shape[-2] - 1
    s_mat = maths.compute_connection_matrix_s(k_max)
    scale = _power_scale(k_max, t_s)
    out = np.empty_like(v)
    out[..., :, 0] = maths.row_times(v[..., ::-1, 0] * scale, s_mat)[..., ::-1]
    out[..., :, 1] = np.sqrt(maths.row_times((v[..., ::-1, 1] * scale) ** 2,
                                             s_mat**2))[..., ::-1]
    return out


def taylor_to_cheby_errors(taylor_error_vec: np.ndarray, t_s: float) -> np.ndarray:
    """Cell widths only (``transforms.py:503-529``)."""
    _check_ts(t_s)
    v = np.asarray(taylor_error_vec, dtype=np.float64)
    k_max = v.shape[-1] - 1
    s_mat = maths.compute_connection_matrix_s(k_max)
    scaled = v[..., ::-1] * _power_scale(k_max, t_s)
    return np.sqrt(maths.row_times(scaled**2, s_mat**2))[..., ::-1]


def cheby_to_taylor(alpha_vec: np.ndarray, t_s: float) -> np.ndarray:
    """Chebyshev to Taylor values (``transforms.py:532-560``): ``(alpha @ R) k! / t_s^k``."""
    _check_ts(t_s)
    a = np.asarray(alpha_vec, dtype=np.float64)
    k_max = a.shape[-1] - 1
    d = maths.row_times(a[..., ::-1], maths.compute_connection_matrix_r(k_max))
    for m in range(k_max + 1):
        d[..., m] = (d[..., m] * fact(m)) / powi(t_s, m)
    return d[..., ::-1]


def cheby_to_taylor_full(alpha_full_vec: np.ndarray, t_s: float) -> np.ndarray:
    """Values and widths (``transforms.py:563-601``)."""
    _check_ts(t_s)
    a = np.asarray(alpha_full_vec, dtype=np.float64)
    k_max = a.shape[-2] - 1
    r_mat = maths.compute_connection_matrix_r(k_max)
    dv = maths.row_times(a[..., ::-1, 0], r_mat)
    var = maths.row_times(a[..., ::-1, 1] ** 2, r_mat**2)
    scale = np.array([fact(k) / powi(t_s, k) for k in range(k_max + 1)])
    out = np.empty_like(a)
    out[..., :, 0] = (dv * scale)[..., ::-1]
    out[..., :, 1] = np.sqrt(var * scale**2)[..., ::-1]
    return out


def _tile_cheby_errors(err: np.ndarray, c_mat: np.ndarray, tiling_strategy: str) -> np.ndarray:
    if tiling_strategy == "conservative":
        return maths.row_times(err, np.abs(c_mat))
    if tiling_strategy == "quadrature":
        return np.sqrt(maths.row_times(err**2, c_mat**2))
    if tiling_strategy == "aggressive":
        return err * np.abs(np.diag(c_mat))
    msg = f"Invalid tiling strategy: {tiling_strategy}"
    raise ValueError(msg)


def shift_cheby_errors(alpha_error_vec: np.ndarray, coord_next: tuple[float, float],
                       coord_cur: tuple[float, float],
                       tiling_strategy: str = "aggressive") -> np.ndarray:
    """Cell widths moved from the domain coord_cur to coord_next (``transforms.py:604-647``)."""
    err = np.asarray(alpha_error_vec, dtype=np.float64)
    c_mat = maths.poly_chebyshev_transform_matrix(err.shape[-1] - 1, coord_cur[0], coord_cur[1],
                                                  coord_next[0], coord_next[1], 1)
    return _tile_cheby_errors(err, c_mat, tiling_strategy)


def shift_cheby_full(alpha_full_vec: np.ndarray, coord_next: tuple[float, float],
                     coord_cur: tuple[float, float],
                     tiling_strategy: str = "aggressive") -> np.ndarray:
    """Values and widths moved to coord_next (``transforms.py:650-700``)."""
    a = np.asarray(alpha_full_vec, dtype=np.float64)
    c_mat = maths.poly_chebyshev_transform_matrix(a.shape[-2] - 1, coord_cur[0], coord_cur[1],
                                                  coord_next[0], coord_next[1], 1)
    out = np.empty_like(a)
    out[..., 0] = maths.row_times(a[..., 0], c_mat)
    out[..., 1] = _tile_cheby_errors(a[..., 1], c_mat, tiling_strategy)
    return out


def cheby_to_taylor_param_shift(alpha_param_vec: np.ndarray, t0: float, ts: float,
                                t_eval: float) -> np.ndarray:
    """Taylor coefficients at t_eval of the series on (t0, ts) (``transforms.py:703-737``)."""
    a = np.asarray(alpha_param_vec, dtype=np.float64)
    c_mat = maths.poly_chebyshev_transform_matrix(a.shape[-1] - 1, t0, ts, t_eval, ts, 1)
    return cheby_to_taylor(maths.row_times(a, c_mat), ts)
