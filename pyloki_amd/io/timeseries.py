"""Time-series container; the part of ``pyloki.io.timeseries.TimeSeries``
(``src/pyloki/io/timeseries.py:16-40``) that the search path reads: ``ts_e`` and ``ts_v``
as float32, the sampling time ``dt``, ``nsamps`` and ``tobs``."""

from __future__ import annotations

import numpy as np


class TimeSeries:
    def __init__(self, ts_e: np.ndarray, ts_v: np.ndarray, dt: float) -> None:
        self._ts_e = np.asarray(ts_e, dtype=np.float32)
        self._ts_v = np.asarray(ts_v, dtype=np.float32)
        self._dt = dt

    @property
    def ts_e(self) -> np.ndarray:
        return self._ts_e

    @property
    def ts_v(self) -> np.ndarray:
        return self._ts_v

    @property
    def dt(self) -> float:
        return self._dt

    @property
    def nsamps(self) -> int:
        return len(self.ts_e)

    @property
    def tobs(self) -> float:
        return self.nsamps * self.dt
