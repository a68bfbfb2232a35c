"""Boxcar matched-filter S/N scoring.

Host side of ``pyloki.detection.scoring`` (``src/pyloki/detection/scoring.py``):
``generate_box_width_trials`` is host planning (``scoring.py:10-39``); the profile
scoring itself (``boxcar_snr_1d/2d/nd``, ``scoring.py:99-165``, and the EP batch scorer
``snr_score_batch_func``, ``scoring.py:249-282``) runs in the HIP library
(``pyl_score_boxcar``). The per-width template constants ``h`` and ``b`` are computed
here in FP64 exactly as the reference computes them and handed to the kernel.
"""

from __future__ import annotations

import math

import numpy as np


def generate_box_width_trials(
    nbins: int,
    ducy_max: float = 0.2,
    wtsp: float = 1.5,
) -> np.ndarray:
    """Geometric boxcar width trials; ``scoring.py:10-39``."""
    wmax = int(max(1, ducy_max * nbins))
    widths = [1]
    while widths[-1] <= wmax:
        next_width = int(max(widths[-1] + 1, int(wtsp * widths[-1])))
        if next_width > wmax:
            break
        widths.append(next_width)
    return np.asarray(widths, dtype=np.int64)


def boxcar_coeffs(nbins: int, widths: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Template height ``h = sqrt((B-w)/(B*w))`` and offset ``b = w*h/(B-w)`` (FP64).

    ``scoring.py:135-137`` and ``:275-277``: the S/N for width ``w`` is
    ``(h + b) * max_i(S[i+w] - S[i]) - b * S[B-1]`` over the circular prefix sum S.
    """
    widths = np.asarray(widths, dtype=np.int64)
    if widths.size == 0:
        msg = "widths must not be empty"
        raise ValueError(msg)
    if np.any(widths < 1) or np.any(widths >= nbins):
        msg = f"widths must lie in [1, nbins-1], got {widths}"
        raise ValueError(msg)
    h = np.empty(widths.size, dtype=np.float64)
    b = np.empty(widths.size, dtype=np.float64)
    for i, w in enumerate(widths.tolist()):
        size_w = nbins - w
        h[i] = math.sqrt(size_w / (nbins * w))
        b[i] = w * h[i] / size_w
    return h, b
