"""Time-series container (``src/pyloki/io/timeseries.py:16-235``): ``ts_e`` and ``ts_v``
as float32, the sampling time ``dt``, ``nsamps``, ``tobs``, ``downsample`` and
``from_tim``.

``from_tim`` reads a SIGPROC ``.tim`` or a PRESTO ``.dat`` file (``io/sigproc.py``), then
removes the red noise and normalises on the GPU, the steps the reference asks sigpyproc
for: ``deredden(method="median", window=filter_window, fast=True)`` and
``normalise(loc_method="mean", scale_method=norm_scale_method)``. The baseline is the
running median in its fast form: medians of consecutive windows of
``int(filter_window / tsamp)`` samples, interpolated linearly between the window centres
(``pyl_ts_deredden_median``); the location is the mean and the scale the standard
deviation, the interquartile range / 1.349 or the median absolute deviation x 1.4826
(``pyl_ts_normalise``).
"""

from __future__ import annotations

import ctypes

import numpy as np

_SCALE = {"std": 0, "iqr": 1, "mad": 2}


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

    def downsample(self, factor: int) -> TimeSeries:
        """Means of ``factor`` consecutive samples (``np_utils.downsample_1d``: the length
        must be a multiple of the factor), the sampling time unchanged as in the
        reference (``timeseries.py:42-45``)."""
        factor = int(factor)
        if factor < 1 or self.nsamps % factor:
            msg = f"cannot downsample {self.nsamps} samples by {factor}"
            raise ValueError(msg)
        e = self.ts_e.reshape(-1, factor).mean(axis=1, dtype=np.float32)
        v = self.ts_v.reshape(-1, factor).mean(axis=1, dtype=np.float32)
        return TimeSeries(e, v, self.dt)

    @classmethod
    def from_tim(cls, timfile: str, tim_type: str = "dat", filter_window: float = 1.0,
                 norm_scale_method: str = "iqr") -> TimeSeries:
        """Load a .tim (SIGPROC) or .dat (PRESTO) series, deredden and normalise it on
        the GPU; ``ts_v`` is all ones (``timeseries.py:197-235``)."""
        from pyloki_amd.io import sigproc

        if tim_type == "dat":
            data, info = sigproc.read_presto_dat(timfile)
        elif tim_type == "tim":
            data, info = sigproc.read_sigproc_tim(timfile)
        else:
            msg = f"Invalid tim type: {tim_type}"
            raise ValueError(msg)
        if norm_scale_method not in _SCALE:
            msg = f"unknown scale method {norm_scale_method!r} (std, iqr or mad)"
            raise ValueError(msg)
        tsamp = float(info["tsamp"])
        signal = condition(data, tsamp, filter_window, norm_scale_method)[0]
        return cls(signal, np.ones_like(signal), tsamp)

    def __str__(self) -> str:
        name = type(self).__name__
        return (f"{name} {{nsamps = {self.nsamps:d}, tsamp = {self.dt:.4e}, "
                f"tobs = {self.tobs:.3f}}}")

    def __repr__(self) -> str:
        return str(self)


def deredden_median(x, window: int):
    """The series minus its median baseline (device tensors): ``(out, block_medians)``."""
    import torch

    from pyloki_amd import _lib
    from pyloki_amd.device import ptr, stream_ptr

    n = int(x.numel())
    window = int(window)
    if window < 1:
        msg = "the filter window is shorter than one sample"
        raise ValueError(msg)
    med = torch.empty(-(-n // window), dtype=torch.float32, device=x.device)
    out = torch.empty_like(x)
    _lib.check(_lib.load().pyl_ts_deredden_median(ptr(x), n, window, ptr(med), ptr(out),
                                                  stream_ptr()))
    return out, med


def normalise(x, scale_method: str = "iqr"):
    """(x - mean) / scale (device tensors): ``(out, mean, scale)``."""
    import torch

    from pyloki_amd import _lib
    from pyloki_amd.device import ptr, stream_ptr

    out = torch.empty_like(x)
    ls = (ctypes.c_double * 2)()
    _lib.check(_lib.load().pyl_ts_normalise(ptr(x), int(x.numel()), _SCALE[scale_method],
                                            ptr(out), ls, stream_ptr()))
    return out, float(ls[0]), float(ls[1])


def condition(data: np.ndarray, tsamp: float, filter_window: float, scale_method: str):
    """``deredden`` then ``normalise`` on the GPU: (FP32 samples, info)."""
    import torch

    from pyloki_amd.device import require_gpu

    dev = require_gpu()
    x = torch.from_numpy(np.ascontiguousarray(data, dtype=np.float32)).to(dev)
    window = int(filter_window / tsamp)
    red, med = deredden_median(x, window)
    out, loc, scale = normalise(red, scale_method)
    return out.cpu().numpy(), {"window": window, "loc": loc, "scale": scale,
                               "nblocks": int(med.numel())}
