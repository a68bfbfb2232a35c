"""Threshold schedules for the EP search."""

from __future__ import annotations

import numpy as np


def bound_scheme(nstages: int, snr_bound: float) -> np.ndarray:
    """``schemes.py:29-50``: thresholds whose squares grow linearly with the stage, up to
    ``snr_bound**2`` at the last segment; FP32, one per stage (level 1 .. nstages)."""
    nsegments = nstages + 1
    thresh_sn2 = (np.arange(1, nsegments + 1) * snr_bound**2 / nsegments).astype(np.float32)
    return np.sqrt(thresh_sn2[1:])
