"""Host-side grid planning math (FP64), the part of the search that is not per-leaf.

Every function restates the pyloki function it names, in the same operation order,
so that grid sizes and grid values are identical to the reference:

- ``poly_taylor_step_f`` / ``poly_taylor_step_d_f``: ``src/pyloki/utils/psr_utils.py:59-113``
- ``range_param_count`` / ``range_param``: ``src/pyloki/utils/psr_utils.py:401-463``
- ``get_phase_idx`` / ``get_phase_idx_int`` (scalar host versions used for planning and
  for small host-side evaluations): ``src/pyloki/utils/psr_utils.py:12-56``

The per-leaf versions of the phase and index math run on the GPU in the HIP kernels.
"""

from __future__ import annotations

import math

import numpy as np

from pyloki_amd.utils.misc import C_VAL, FLOAT_EPSILON


def powi(a: float, b: int) -> float:
    """``a**b`` for an integer ``b`` by binary powering.

    The reference evaluates ``x**k`` for integer arrays ``k`` in numba fastmath code
    (``psr_utils.py:89``, ``transforms.py:43``); LLVM's fast-math simplification rewrites
    ``pow(x, sitofp(k))`` to ``llvm.powi``, whose runtime routine is this multiply chain.
    """
    recip = b < 0
    b = -b if recip else b
    r = 1.0
    while True:
        if b & 1:
            r *= a
        b //= 2
        if b == 0:
            break
        a *= a
    return 1.0 / r if recip else r


def fact(n: int) -> float:
    """Factorial as a float, like ``maths.fact`` (a float64 table, ``maths.py:11-33``)."""
    return float(math.factorial(int(n)))


def poly_taylor_step_f(
    nparams: int,
    tobs: float,
    nbins: int,
    eta: float,
    t_ref: float = 0.0,
    use_cheby: bool = True,
) -> np.ndarray:
    """Grid step for {f_k, ..., f_0}; ``psr_utils.py:59-96``.

    ``dphi * (k+1)! / (tobs - t_ref)^(k+1)``, times ``2^k`` with Chebyshev coarsening,
    returned highest derivative first.
    """
    dphi = eta / nbins
    out = np.empty(nparams, dtype=np.float64)
    base = tobs - t_ref
    for k in range(nparams):
        val = dphi * fact(k + 1) / powi(base, k + 1)
        if use_cheby:
            val = (2**k) * val
        out[k] = val
    return out[::-1].copy()


def poly_taylor_step_d_f(
    nparams: int,
    tobs: float,
    nbins: int,
    eta: float,
    f_max: float,
    t_ref: float = 0.0,
    use_cheby: bool = True,
) -> np.ndarray:
    """Grid for {d_k, ..., d_2, f}; ``psr_utils.py:99-113``."""
    dparams_f = poly_taylor_step_f(nparams, tobs, nbins, eta, t_ref, use_cheby)
    dparams = np.zeros(nparams, dtype=np.float64)
    dparams[:-1] = dparams_f[:-1] * C_VAL / f_max
    dparams[-1] = dparams_f[-1]
    return dparams


def poly_taylor_step_d(
    poly_order: int,
    tobs: float,
    nbins: int,
    eta: float,
    f_max: float,
    t_ref: float = 0.0,
    use_cheby: bool = True,
) -> np.ndarray:
    """Grid for {d_k_max, ..., d_1}; ``psr_utils.py:116-128``."""
    dparams_f = poly_taylor_step_f(poly_order, tobs, nbins, eta, t_ref, use_cheby)
    return dparams_f * C_VAL / f_max


def range_param_count(vmin: float, vmax: float, dv: float) -> int:
    """Number of cell centres of ``range_param``; ``psr_utils.py:401-425``."""
    if not (vmin < vmax and dv > 0.0):
        msg = "Invalid input: ensure vmin < vmax and dv > 0.0."
        raise ValueError(msg)
    if dv >= (vmax - vmin):
        return 1
    return math.ceil((vmax - vmin) / dv)


def linspace(start: float, stop: float, num: int) -> np.ndarray:
    """``np.linspace`` as numba and NumPy evaluate it: ``start + i*step``, last = stop."""
    out = np.empty(num, dtype=np.float64)
    if num == 0:
        return out
    div = num - 1
    if div > 0:
        step = (stop - start) / div
        out[:] = start + np.arange(num, dtype=np.float64) * step
        out[-1] = stop
    else:
        out[0] = start
    return out


def range_param(vmin: float, vmax: float, dv: float) -> np.ndarray:
    """Cell centres that tile [vmin, vmax]; ``psr_utils.py:428-463``."""
    if not (vmin < vmax and dv > 0):
        msg = "Invalid input: ensure vmin < vmax and dv > 0."
        raise ValueError(msg)
    if dv >= (vmax - vmin):
        return np.array([(vmax + vmin) / 2.0])
    npoints = math.ceil((vmax - vmin) / dv)
    dv_actual = (vmax - vmin) / npoints
    return linspace(vmin + (dv_actual / 2.0), vmax - (dv_actual / 2.0), npoints)


def get_phase_idx(proper_time: float, freq: float, nbins: int, delay: float) -> float:
    """Fractional phase bin in [0, nbins); ``psr_utils.py:12-46``."""
    if freq <= 0:
        msg = "Frequency must be positive."
        raise ValueError(msg)
    if nbins <= 0:
        msg = "Number of bins must be positive."
        raise ValueError(msg)
    phase = (proper_time - delay) * freq
    phase -= math.floor(phase)
    iphase = phase * nbins
    if iphase >= nbins:
        iphase = 0.0
    return iphase


def get_phase_idx_int(proper_time: float, freq: float, nbins: int, delay: float) -> int:
    """Integer phase bin, half-up rounding in FP32; ``psr_utils.py:49-56``."""
    shifts = get_phase_idx(proper_time, freq, nbins, delay)
    ibin = int(np.float32(shifts) + np.float32(0.5))
    if ibin == nbins:
        ibin = 0
    return ibin


def get_nearest_indices_analytical(
    param_set: np.ndarray,
    param_grid_count: np.ndarray,
    param_limits: np.ndarray,
) -> np.ndarray:
    """Grid index by truncation with a 1e-6 guard; ``psr_utils.py:466-514``."""
    nparams = len(param_set)
    pindex = np.zeros(nparams, dtype=np.int64)
    for ip in range(nparams):
        n_grid = int(param_grid_count[ip])
        if n_grid <= 1:
            continue
        vmin, vmax = float(param_limits[ip][0]), float(param_limits[ip][1])
        raw_idx = n_grid * (float(param_set[ip]) - vmin) / (vmax - vmin)
        idx = int(raw_idx + FLOAT_EPSILON)
        pindex[ip] = min(max(idx, 0), n_grid - 1)
    return pindex
