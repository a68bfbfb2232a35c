"""Seeded synthetic pulsar time series (input generation for tests and benchmarks).

A seeded counterpart of ``PulseSignalConfig.generate`` (``src/pyloki/simulation/pulse.py``,
whose random generators are unseeded): N(0, 1) white noise plus a Gaussian pulse whose
arrival times follow the Taylor delay model of ``DerivativeModulating``
(``modulate.py:24-50``: ``t_proper = t - delay(t - t_ref)/c``) or a circular orbit.
The pulse amplitude is set analytically so that the optimal matched-filter S/N of the
whole series equals ``snr``: ``A = snr / sqrt(sum(template**2))``. Everything is host
NumPy (float64 phases, float32 output).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from pyloki_amd.io.timeseries import TimeSeries
from pyloki_amd.utils.misc import C_VAL


@dataclass(kw_only=True)
class PulseSignalConfig:
    period: float
    dt: float
    nsamps: int = 2**21
    snr: float = 100.0
    ducy: float = 0.1
    mod_kwargs: dict = field(default_factory=dict)

    @property
    def freq(self) -> float:
        return 1.0 / self.period

    @property
    def tobs(self) -> float:
        return self.nsamps * self.dt

    @property
    def fold_bins(self) -> int:
        return int(max(16, min(256, round(1.0 / self.ducy * 8))))

    def proper_time(self, t: np.ndarray) -> np.ndarray:
        kw = dict(self.mod_kwargs)
        t_ref = kw.pop("t_ref", self.tobs / 2)
        if "p_orb" in kw:
            # Circular orbit: delay = x * c * sin(omega * (t - t_ref) + phi) (light-seconds).
            omega = 2.0 * np.pi / kw["p_orb"]
            delay_s = kw["x_orb"] * np.sin(omega * (t - t_ref) + kw.get("phi", 0.0))
            return t - delay_s
        x = t - t_ref
        delay = kw.get("shift", 0.0) + x * (
            kw.get("vel", 0.0)
            + x * (kw.get("acc", 0.0) * 0.5
                   + x * (kw.get("jerk", 0.0) / 6.0 + x * kw.get("snap", 0.0) / 24.0))
        )
        return t - delay / C_VAL

    def template(self) -> np.ndarray:
        t = np.arange(self.nsamps, dtype=np.float64) * self.dt
        tp = self.proper_time(t)
        phase = np.mod(tp * self.freq, 1.0)
        sigma = self.ducy / (2.0 * np.sqrt(2.0 * np.log(10.0)))
        d = np.minimum(np.abs(phase - 0.5), 1.0 - np.abs(phase - 0.5))
        return np.exp(-0.5 * (d / sigma) ** 2)

    def generate(self, shape: str = "gaussian", seed: int = 0) -> TimeSeries:
        if shape != "gaussian":
            msg = f"Unknown shape: {shape}"
            raise ValueError(msg)
        rng = np.random.default_rng(seed)
        noise = rng.standard_normal(self.nsamps, dtype=np.float32)
        tmpl = self.template()
        tmpl -= tmpl.mean()
        amp = self.snr / np.sqrt(np.sum(tmpl * tmpl)) if self.snr > 0 else 0.0
        ts_e = (noise + (amp * tmpl).astype(np.float32)).astype(np.float32)
        ts_v = np.ones(self.nsamps, dtype=np.float32)
        return TimeSeries(ts_e, ts_v, self.dt)


def generate_folded_profile(nbins: int = 100, ducy: float = 0.1, center: float = 0.5
                            ) -> np.ndarray:
    """Unit-energy Gaussian folded profile (``pulse.py:512-523``): full width at a tenth
    of the peak ``ducy`` of a turn, peak at phase ``center``, wrapped; FP64."""
    phase = np.linspace(0, 1, nbins, endpoint=False)
    sigma = ducy / (2 * np.sqrt(2 * np.log(10)))
    wrapped = np.mod(phase - center + 0.5, 1) - 0.5
    profile = np.exp(-(wrapped**2) / (2 * sigma**2))
    profile = profile / profile.max()
    return profile / np.sqrt(np.sum(profile**2))
