"""Periodogram of an FFA search: the read-out API the notebooks use
(``src/pyloki/periodogram.py:24-87``), on plain NumPy arrays instead of xarray.
Dimensions are the search parameters in ``crackle..accel, freq`` order, then ``width``."""

from __future__ import annotations

from typing import ClassVar

import numpy as np


def find_nearest_sorted_idx(array: np.ndarray, value: float, rtol: float = 1e-5,
                            atol: float = 1e-8) -> int:
    """Index of the closest value in a sorted array, ties to the smaller value
    (``np_utils.py:13-55``)."""
    n = len(array)
    if n == 0:
        msg = "Array must not be empty"
        raise ValueError(msg)
    idx = int(np.searchsorted(array, value, side="left"))
    if idx > 0:
        diff_prev = abs(value - array[idx - 1])
        diff_curr = abs(value - array[idx]) if idx < n else np.inf
        if diff_prev <= diff_curr * (1 + rtol) + atol:
            return idx - 1
    return min(idx, n - 1)


class Periodogram:
    REQUIRED_PARAMS: ClassVar[tuple[str, str]] = ("freq", "width")
    OPTIONAL_PARAMS: ClassVar[tuple[str, ...]] = ("crackle", "snap", "jerk", "accel")

    def __init__(self, params: dict[str, np.ndarray], snrs: np.ndarray, tobs: float) -> None:
        self.params = params
        self.snrs = np.asarray(snrs)
        self.tobs = tobs
        for name in self.REQUIRED_PARAMS:
            if name not in params:
                msg = f"missing required parameter: {name}"
                raise ValueError(msg)
        shape = tuple(len(params[n]) for n in self.dims)
        if self.snrs.shape != shape:
            msg = f"snrs shape {self.snrs.shape} does not match params {shape}"
            raise ValueError(msg)

    @property
    def valid_params(self) -> tuple[str, ...]:
        return self.OPTIONAL_PARAMS + self.REQUIRED_PARAMS

    @property
    def param_names(self) -> list[str]:
        return [p for p in self.valid_params if p in self.params]

    @property
    def dims(self) -> list[str]:
        names = [p for p in self.OPTIONAL_PARAMS + ("freq",) if p in self.params]
        return [*names, "width"]

    @property
    def ndim(self) -> int:
        return len(self.param_names)

    def find_best_indices(self) -> tuple[int, ...]:
        flat_index = np.argmax(self.snrs)
        return tuple(int(i) for i in np.unravel_index(int(flat_index), self.snrs.shape))

    def find_best_params(self) -> dict[str, float]:
        best = self.find_best_indices()
        out = {"snr": float(self.snrs[best])}
        for name, idx in zip(self.dims, best):
            out[name] = self.params[name][idx]
        return out

    def get_indices_summary(self, true_values: dict[str, float] | None = None) -> str:
        best = self.find_best_indices()
        summary = []
        if true_values:
            true_idx = [find_nearest_sorted_idx(self.params[n], true_values[n])
                        for n in self.param_names if n in true_values]
            summary.append(f"True param indices: {tuple(true_idx)}")
        summary.append(f"Best param indices: {best}")
        return "\n".join(summary)

    def get_summary(self) -> str:
        best = self.find_best_params()
        summary = [f"Best S/N: {best['snr']:.2f}", f"Best Period: {1 / best['freq']}"]
        summary += [f"Best {n}: {best[n]}" for n in self.param_names if n != "snr"]
        return "\n".join(summary)
