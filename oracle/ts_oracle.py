"""CPU oracle of the time-series conditioning -- TEST INFRASTRUCTURE ONLY.

Restates the steps ``TimeSeries.from_tim`` (``io/timeseries.py:197-235``) takes from
sigpyproc: the median dereddening in its fast form and the normalisation. sigpyproc is an
unpinned git dependency of the reference and is absent here, so these follow the
product's documented definitions (include/pyloki_hip.h, "Time-series conditioning"):
the quantile of m values at q is numpy's "linear" one from the exact order statistics at
floor and ceil of q (m - 1), interpolated as a + (b - a) f in FP64; the baseline
interpolates the block medians between block centres in FP64. Parity with sigpyproc is
unpinned; the definitions are pinned by tests/test_ts_oracle.py (numpy's percentile and
median, exact order statistics on integer data, a ramp's baseline equal to the ramp,
closed-form locations and scales).
"""

from __future__ import annotations

import numpy as np

IQR_NORMAL = 1.3489795003921634
MAD_NORMAL = 1.482602218505602


def quantile(x: np.ndarray, q: float) -> np.float32:
    x = np.sort(np.asarray(x, dtype=np.float32))
    m = len(x)
    p = q * (m - 1)
    lo = int(np.floor(p))
    hi = min(lo + 1, m - 1)
    f = p - np.floor(p)
    a, b = float(x[lo]), float(x[hi])
    return np.float32(a + (b - a) * f)


def segment_quantiles(x: np.ndarray, seg_len: int, qs, center=None) -> np.ndarray:
    n = len(x)
    nseg = -(-n // seg_len)
    out = np.empty((nseg, len(qs)), dtype=np.float32)
    for s in range(nseg):
        seg = np.asarray(x[s * seg_len:(s + 1) * seg_len], dtype=np.float32)
        if center is not None:
            seg = np.abs(seg - np.float32(center[s])).astype(np.float32)
        for t, q in enumerate(qs):
            out[s, t] = quantile(seg, q)
    return out


def deredden_median(x: np.ndarray, w: int) -> tuple[np.ndarray, np.ndarray]:
    x = np.asarray(x, dtype=np.float32)
    n = len(x)
    med = segment_quantiles(x, w, [0.5])[:, 0]
    nblk = len(med)
    sizes = np.minimum(w, n - np.arange(nblk) * w)
    centres = np.arange(nblk) * w + 0.5 * (sizes - 1)
    i = np.arange(n)
    j = i // w
    right = i >= centres[j]
    j0 = np.where(right, j, np.maximum(j - 1, 0))
    j1 = np.where(right, np.minimum(j + 1, nblk - 1), j)
    c0, c1 = centres[j0], centres[j1]
    m0, m1 = med[j0].astype(np.float64), med[j1].astype(np.float64)
    same = j0 == j1
    t = np.where(same, 0.0, (i - c0) / np.where(same, 1.0, c1 - c0))
    base = np.where(same, m0, m0 + (m1 - m0) * t)
    return (x.astype(np.float64) - base).astype(np.float32), med


def normalise(x: np.ndarray, method: str) -> tuple[np.ndarray, float, float]:
    x = np.asarray(x, dtype=np.float32)
    x64 = x.astype(np.float64)
    loc = float(np.sum(x64) / len(x))
    if method == "std":
        scale = float(np.sqrt(np.sum((x64 - loc) ** 2) / len(x)))
    elif method == "iqr":
        scale = (float(quantile(x, 0.75)) - float(quantile(x, 0.25))) / IQR_NORMAL
    elif method == "mad":
        med = quantile(x, 0.5)
        scale = float(quantile(np.abs(x - med).astype(np.float32), 0.5)) * MAD_NORMAL
    else:
        raise ValueError(method)
    return ((x64 - loc) / scale).astype(np.float32), loc, scale
