"""Pulsar signal simulation (``src/pyloki/simulation/pulse.py``).

``PulseSignalConfig`` keeps the reference's fields and ``generate(shape, phi0, max_iter,
tol)`` (``pulse.py:310-406``). Its sample-level work runs on the GPU in one launch
(``pyl_sim_series``): the modulated proper time, the pulse template integrated from the
shape's CDF table, N(0, 1) noise from the counter-based generator, and the two calibration
folds at the true period. The calibration itself works on those folded profiles (one
profile of ``int(period / dt)`` bins, host NumPy in the reference's dtypes), and a second
launch (``pyl_sim_combine``) forms the series.

Differences from the reference, by design:
- the noise is seeded (``seed=``) and reproducible: the reference draws an unseeded numpy
  Generator, so no two of its runs agree and nothing can be compared bit for bit with it;
- the calibration folds are exact fixed-point sums, not FP32 running sums in sample order.

``generate_analytic(seed)`` is the seeded host generator the tests and benchmarks use for
inputs on machines without a GPU: numpy noise plus a Gaussian pulse sampled at each
sample's phase with the amplitude fixed analytically (``snr / sqrt(sum(template**2))``).
"""

from __future__ import annotations

import ctypes
import dataclasses
from dataclasses import dataclass, field

import numpy as np

from pyloki_amd.io.timeseries import TimeSeries
from pyloki_amd.simulation.modulate import type_to_mods


def build_cdf_lut(shape: str, width: float, pos: float, ngrid: int = 4096) -> np.ndarray:
    """``pulse.py:13-36``: the pulse shape's CDF on ``ngrid`` phases with the first value
    repeated at the end (for the interpolation's wrap), FP32."""
    from scipy import stats

    phase = np.linspace(0.0, 1.0, ngrid, endpoint=False)
    if shape == "boxcar":
        rv = stats.uniform(loc=pos, scale=width)
    elif shape == "gaussian":
        rv = stats.norm(loc=pos, scale=width / (2 * np.sqrt(2 * np.log(10))))
    elif shape == "von_mises":
        kappa = np.log(2.0) / (2.0 * np.sin(np.pi * width / 2.0) ** 2)
        rv = stats.vonmises(loc=pos, kappa=kappa)
    else:
        msg = f"Unknown shape: {shape}"
        raise ValueError(msg)
    cdf = rv.cdf(phase).astype(np.float64)
    return np.concatenate([cdf, cdf[:1]]).astype(np.float32)


def boxcar_snr_1d(norm_data: np.ndarray, widths: np.ndarray, stdnoise: float = 1.0
                  ) -> np.ndarray:
    """Boxcar S/N of one normalised profile (``scoring.py:105-140``): FP32 inclusive
    circular prefix sum, FP64 height and offset per width, FP32 result."""
    x = np.asarray(norm_data, dtype=np.float32)
    size = len(x)
    widths = np.asarray(widths, dtype=np.int64)
    wmax = int(widths.max())
    ps = np.cumsum(x, dtype=np.float32)
    last = ps[size - 1]
    reps = [ps]
    for i in range(1, (size + wmax) // size + 1):
        reps.append(ps + np.float32(i) * last)
    ext = np.concatenate(reps)[: size + wmax]
    snr = np.empty(len(widths), dtype=np.float32)
    for iw, w in enumerate(widths):
        size_w = size - int(w)
        h = np.sqrt(size_w / (size * int(w)))
        b = int(w) * h / size_w
        dmax = float(np.max(ext[w:w + size] - ext[:size]))
        snr[iw] = ((h + b) * dmax - b * float(last)) * (1.0 / stdnoise)
    return snr


def calibrate_scale_on_folds(noise_vals: np.ndarray, noise_var: np.ndarray,
                             tmpl_vals: np.ndarray, snr_target: float, initial_scale: float,
                             max_iter: int = 10, tol: float = 1e-3, damping: float = 0.8
                             ) -> float:
    """``pulse.py:69-111``: damped Newton steps on the template scale until the folded
    profile's best boxcar S/N is within ``tol`` of the target (fold() is linear, so the
    profile of noise + s * template is folded noise + s * folded template)."""
    nbins = noise_vals.shape[0]
    widths = np.arange(1, nbins // 2)
    sd = np.sqrt(noise_var.astype(np.float32))
    tmpl_norm = (tmpl_vals.astype(np.float32) / sd).astype(np.float32)
    noise_norm = (noise_vals.astype(np.float32) / sd).astype(np.float32)
    snr_template = float(np.max(boxcar_snr_1d(tmpl_norm, widths, 1.0)))
    scale = float(initial_scale)
    for _ in range(max_iter):
        combined = (noise_norm.astype(np.float64) + scale * tmpl_norm.astype(np.float64)
                    ).astype(np.float32)
        diff = float(np.max(boxcar_snr_1d(combined, widths, 1.0))) - snr_target
        if diff * diff < tol * tol:
            break
        scale -= damping * diff / snr_template
    return scale


def off_pulse_mask(nbins: int, phi0: float, ducy: float) -> np.ndarray:
    """``pulse.py:370-376``: bins outside the pulse's ``round(ducy nbins)`` bins around
    ``round(phi0 nbins)``."""
    center = round(phi0 * nbins) % nbins
    width = max(1, round(ducy * nbins))
    on = (center - (width // 2) + np.arange(width)) % nbins
    mask = np.ones(nbins, dtype=bool)
    mask[on] = False
    return mask


@dataclass(kw_only=True)
class PulseSignalConfig:
    """A pulsar signal generator configuration (``pulse.py:114-177``)."""

    period: float
    dt: float
    nsamps: int = 2**21
    snr: float = 100
    ducy: float = 0.1
    ds: float = 1.0
    mod_type: str = "derivative"
    mod_kwargs: dict = field(default_factory=dict)
    mod_tref: float | None = None

    def __post_init__(self) -> None:
        if self.mod_type not in type_to_mods:
            msg = f"Unknown modulation type: {self.mod_type}"
            raise ValueError(msg)
        self._mod_func = type_to_mods[self.mod_type](**self.mod_kwargs)
        if self.ducy <= 0 or self.ducy >= 1:
            msg = f"Duty cycle ({self.ducy}) should be in (0, 1)"
            raise ValueError(msg)
        if self.mod_tref is None:
            self.mod_tref = self.tobs / 2

    @property
    def freq(self) -> float:
        return 1 / self.period

    @property
    def tobs(self) -> float:
        return self.nsamps * self.dt

    @property
    def tol_bins(self) -> float:
        return self.ducy * self.period / self.dt

    @property
    def fold_bins(self) -> int:
        return int(self.period / self.dt / self.ds)

    @property
    def fold_bins_ideal(self) -> int:
        return int(self.period / self.dt)

    @property
    def mod_func(self):
        return self._mod_func

    def time_array(self) -> np.ndarray:
        """Sample times ``i * dt`` (the values ``numpy.arange(0, tobs, dt)`` forms)."""
        return np.arange(self.nsamps, dtype=np.float64) * self.dt

    @property
    def proper_time(self) -> np.ndarray:
        return self.mod_func.generate(self.time_array(), self.mod_tref)

    def get_updated(self, update_dict: dict | None) -> PulseSignalConfig:
        new = {f.name: getattr(self, f.name) for f in dataclasses.fields(self)}
        if update_dict is not None:
            new.update({k: v for k, v in update_dict.items() if k in new})
        return PulseSignalConfig(**new)

    # -- GPU generation ---------------------------------------------------------------
    def simulate_device(self, shape: str = "gaussian", phi0: float = 0.5, max_iter: int = 5,
                        tol: float = 1e-2, *, seed: int | None = None, stream_id: int = 0):
        """The series on the device: ``(ts_e, ts_v, info)`` with FP32 device tensors and
        the calibration's figures (noise and signal scales, folded S/N of the template)."""
        import torch

        from pyloki_amd import _lib
        from pyloki_amd.device import ptr, require_gpu, stream_ptr

        dev = require_gpu()
        lib = _lib.load()
        if seed is None:
            seed = int(np.random.default_rng().integers(0, 2**63))
        ngrid = max(4096, int(100 / self.ducy))
        lut = torch.from_numpy(build_cdf_lut(shape, self.ducy, phi0, ngrid=ngrid)).to(dev)
        nbins = self.fold_bins_ideal
        if nbins < 1:
            msg = f"period ({self.period}) is shorter than one sample ({self.dt})"
            raise ValueError(msg)
        n = int(self.nsamps)
        noise = torch.empty(n, dtype=torch.float32, device=dev)
        tmpl = torch.empty(n, dtype=torch.float32, device=dev)
        folds = torch.empty(3 * nbins, dtype=torch.int64, device=dev)
        mod = self.mod_func.to_device(self.mod_tref)
        _lib.check(lib.pyl_sim_series(ctypes.byref(mod), n, float(self.dt), float(self.period),
                                      ptr(lut), ngrid, ctypes.c_uint64(seed),
                                      ctypes.c_uint64(stream_id), nbins, ptr(noise), ptr(tmpl),
                                      ptr(folds), stream_ptr()))
        acc = folds.cpu().numpy().reshape(3, nbins)
        # Fixed-point sums (x 2^32) as the FP32 folded profiles (pulse.py:348-366).
        noise_vals = (acc[0].astype(np.float64) * 2.0**-32).astype(np.float32)
        tmpl_vals = (acc[1].astype(np.float64) * 2.0**-32).astype(np.float32)
        noise_var = acc[2].astype(np.float32)
        ns, cs, snr_tmpl = calibration_scales(noise_vals, noise_var, tmpl_vals, self.snr,
                                              phi0, self.ducy, max_iter, tol)
        ts_e = torch.empty(n, dtype=torch.float32, device=dev)
        ts_v = torch.empty(n, dtype=torch.float32, device=dev)
        _lib.check(lib.pyl_sim_combine(ptr(noise), ptr(tmpl), n, ctypes.c_float(ns),
                                       ctypes.c_float(cs), ctypes.c_float(ns**2), ptr(ts_e),
                                       ptr(ts_v), stream_ptr()))
        info = {"seed": seed, "noise_scale": float(ns), "signal_scale": float(cs),
                "snr_template": float(snr_tmpl), "nbins": nbins,
                "folded": {"noise": noise_vals, "var": noise_var, "template": tmpl_vals}}
        return ts_e, ts_v, info

    def generate(self, shape: str = "gaussian", phi0: float = 0.5, max_iter: int = 5,
                 tol: float = 1e-2, *, seed: int | None = None) -> TimeSeries:
        """``pulse.py:310-406`` on the GPU: a series whose fold at the true ephemeris has
        the configured boxcar S/N. ``seed=None`` draws a fresh seed (the reference's
        behaviour); a seed makes the series reproducible."""
        ts_e, ts_v, _ = self.simulate_device(shape, phi0, max_iter, tol, seed=seed)
        return TimeSeries(ts_e.cpu().numpy(), ts_v.cpu().numpy(), self.dt)

    # -- host generator for test inputs -----------------------------------------------
    def template_analytic(self) -> np.ndarray:
        """Gaussian pulse (FWTM ``ducy``, peak at phase 0.5) sampled at each sample's
        proper-time phase, FP64."""
        phase = np.mod(self.proper_time * self.freq, 1.0)
        sigma = self.ducy / (2.0 * np.sqrt(2.0 * np.log(10.0)))
        d = np.minimum(np.abs(phase - 0.5), 1.0 - np.abs(phase - 0.5))
        return np.exp(-0.5 * (d / sigma) ** 2)

    def generate_analytic(self, seed: int = 0) -> TimeSeries:
        """Seeded host generator (NumPy) for test and benchmark inputs: N(0, 1) noise plus
        the mean-free analytic template scaled to the whole series' matched-filter S/N."""
        rng = np.random.default_rng(seed)
        noise = rng.standard_normal(self.nsamps, dtype=np.float32)
        tmpl = self.template_analytic()
        tmpl -= tmpl.mean()
        amp = self.snr / np.sqrt(np.sum(tmpl * tmpl)) if self.snr > 0 else 0.0
        ts_e = (noise + (amp * tmpl).astype(np.float32)).astype(np.float32)
        return TimeSeries(ts_e, np.ones(self.nsamps, dtype=np.float32), self.dt)


def calibration_scales(noise_vals: np.ndarray, noise_var: np.ndarray, tmpl_vals: np.ndarray,
                       snr: float, phi0: float, ducy: float, max_iter: int, tol: float
                       ) -> tuple[np.float32, float, np.float32]:
    """``pulse.py:368-400`` on the folded profiles: the noise scale from the off-pulse
    bins (FP32), the first scale guess, and the Newton-calibrated signal scale."""
    nbins = len(noise_vals)
    mask = off_pulse_mask(nbins, phi0, ducy)
    noise_scale = np.float32(1.0) / np.std(noise_vals[mask])
    nv_scaled = noise_vals * noise_scale
    var_scaled = noise_var * noise_scale**2
    tmpl_norm = tmpl_vals / np.sqrt(var_scaled)
    snr_template = np.max(boxcar_snr_1d(tmpl_norm, np.arange(1, nbins // 2), 1))
    initial = (np.float32(snr) / snr_template) * noise_scale
    scale = calibrate_scale_on_folds(nv_scaled.astype(np.float32), var_scaled.astype(np.float32),
                                     tmpl_vals.astype(np.float32), snr, float(initial),
                                     max_iter, tol, damping=0.8)
    return noise_scale, scale, snr_template


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
