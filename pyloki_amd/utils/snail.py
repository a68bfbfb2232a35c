"""Middle-out segment scheme of the EP search (``src/pyloki/utils/snail.py:18-390``).

Segments are visited in order of distance from the reference segment (ties: lower index
first); coordinates are ``(reference time, half-span)`` of the segments visited so far.
Only the stride-1 scheme that ``Pruning`` uses is restated.
"""

from __future__ import annotations

import numpy as np


class MiddleOutScheme:
    def __init__(self, nsegments: int, ref_idx: int, tsegment: float, stride: int = 1) -> None:
        if nsegments <= 0:
            msg = f"nsegments must be greater than 0, got {nsegments}."
            raise ValueError(msg)
        if ref_idx < 0 or ref_idx >= nsegments:
            msg = f"ref_idx must be in [0, {nsegments}), got {ref_idx}."
            raise ValueError(msg)
        if tsegment <= 0:
            msg = f"tsegment must be greater than 0, got {tsegment}."
            raise ValueError(msg)
        if stride != 1:
            msg = "only stride 1 is supported"
            raise ValueError(msg)
        self.nsegments = nsegments
        self.ref_idx = ref_idx
        self.tsegment = tsegment
        self.stride = stride
        self.data = np.argsort(np.abs(np.arange(nsegments) - ref_idx), kind="mergesort")

    def _check(self, level: int) -> None:
        if level < 0 or level >= self.nsegments:
            msg = f"level must be in [0, {self.nsegments - 1}], got {level}."
            raise ValueError(msg)

    def get_segment_idx(self, level: int) -> int:
        self._check(level)
        return int(self.data[level])

    def get_coord(self, level: int) -> tuple[float, float]:
        self._check(level)
        till = self.data[: level + 1]
        ref = (int(till.min()) + int(till.max()) + 1) / 2
        scale = ref - int(till.min())
        return ref * self.tsegment, scale * self.tsegment

    def get_segment_coord(self, level: int) -> tuple[float, float]:
        self._check(level)
        return (self.get_segment_idx(level) + 0.5) * self.tsegment, 0.5 * self.tsegment

    def get_segment_coords_so_far(self, level: int) -> tuple[np.ndarray, np.ndarray]:
        self._check(level)
        idx = np.array([self.get_segment_idx(i) for i in range(level + 1)], dtype=np.int64)
        res = np.zeros((level + 1, 2), dtype=np.float64)
        for i in range(level + 1):
            res[i] = self.get_segment_coord(i)
        return idx, res

    def get_current_coord(self, level: int, moving_grid: bool = True) -> tuple[float, float]:
        self._check(level)
        if level == 0:
            return self.get_coord(level)
        if moving_grid:
            return self.get_coord(level - 1)[0], self.get_coord(level)[1]
        t0_init, scale_init = self.get_coord(0)
        t0_next, scale_next = self.get_coord(level)
        scale_cur = max(abs(t0_next - scale_next - t0_init), abs(t0_next + scale_next - t0_init))
        return t0_init, scale_cur - scale_init

    def get_previous_coord(self, level: int, moving_grid: bool = True) -> tuple[float, float]:
        return self.get_current_coord(level - 1, moving_grid) if not moving_grid and level - 1 > 0 \
            else self.get_coord(level - 1)

    def get_report_coord(self, moving_grid: bool = True) -> tuple[float, float]:
        level = self.nsegments - 1
        if moving_grid:
            return self.get_coord(level)
        t0_init, scale_init = self.get_coord(0)
        t0_next, scale_next = self.get_coord(level)
        scale_cur = max(abs(t0_next - scale_next - t0_init), abs(t0_next + scale_next - t0_init))
        return t0_next, scale_cur - scale_init
