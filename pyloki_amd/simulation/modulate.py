"""Time-modulating models: proper time of a barycentric time array
(``src/pyloki/simulation/modulate.py``).

Each model keeps the reference's class name, keyword fields and ``generate(t_arr, t_ref)``
(host NumPy, FP64, the reference's operation order), and packs itself into the C ABI's
``pyl_modulation`` for the GPU signal simulation (``pyl_sim_series``), which evaluates the
same formula per sample on the device.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from pyloki_amd.utils.misc import C_VAL

PYL_MOD_DERIVATIVE = 0
PYL_MOD_DERIVATIVE_SERIES = 1
PYL_MOD_CIRCULAR = 2
PYL_MOD_CIRCULAR_T0 = 3
PYL_MOD_KEPLERIAN = 4
PYL_MOD_MAX_COEFFS = 24


class PylModulation(ctypes.Structure):
    """ctypes mirror of ``pyl_modulation`` (include/pyloki_hip.h)."""

    _fields_ = [("kind", ctypes.c_int32), ("n_coeffs", ctypes.c_int32),
                ("t_ref", ctypes.c_double), ("p", ctypes.c_double * 6),
                ("coeffs", ctypes.c_double * PYL_MOD_MAX_COEFFS)]


def _packed(kind: int, t_ref: float, p=(), coeffs=()) -> PylModulation:
    m = PylModulation()
    m.kind = kind
    m.t_ref = float(t_ref)
    for i, v in enumerate(p):
        m.p[i] = float(v)
    m.n_coeffs = len(coeffs)
    for i, v in enumerate(coeffs):
        m.coeffs[i] = float(v)
    return m


def _fact(k: np.ndarray) -> np.ndarray:
    out = np.ones(len(k), dtype=np.float64)
    for i, n in enumerate(k):
        for j in range(2, int(n) + 1):
            out[i] *= j
    return out


@dataclass(kw_only=True)
class DerivativeModulating:
    """Taylor delay around t_ref up to snap (``modulate.py:23-54``); coefficients are
    displacement derivatives in metres."""

    shift: float = 0.0
    vel: float = 0.0
    acc: float = 0.0
    jerk: float = 0.0
    snap: float = 0.0

    def generate(self, t_arr: np.ndarray, t_ref: float = 0.0) -> np.ndarray:
        dt = t_arr - t_ref
        delay = self.shift + dt * (
            self.vel
            + dt * (self.acc * 0.5 + dt * (self.jerk * (1.0 / 6.0) + dt * self.snap * (1.0 / 24.0)))
        )
        return t_arr - delay / C_VAL

    def to_device(self, t_ref: float) -> PylModulation:
        return _packed(PYL_MOD_DERIVATIVE, t_ref,
                       (self.shift, self.vel, self.acc, self.jerk, self.snap))

    def to_circular(self) -> dict[str, float]:
        """``modulate.py:56-74``: (p_orb, psi, x_orb) of the circular orbit with these
        derivatives at t_ref."""
        if self.acc == 0.0 and self.snap == 0.0:
            msg = "Degenerate phase: cannot recover omega from (d2,d3,d4)."
            raise ValueError(msg)
        if self.acc * self.snap >= 0.0:
            msg = "Incompatible with circular orbit: require d2*d4<0."
            raise ValueError(msg)
        omega = np.sqrt(-self.snap / self.acc)
        x_sin = -self.acc / (C_VAL * omega**2)
        x_cos = -self.jerk / (C_VAL * omega**3)
        return {"p_orb": 2.0 * np.pi / omega, "psi": np.atan2(x_sin, x_cos),
                "x_orb": np.hypot(x_sin, x_cos)}


@dataclass(kw_only=True)
class DerivativeSeriesModulating:
    """Taylor delay of any order (``modulate.py:77-119``): ``coeffs[k]`` = d^k(t_ref)."""

    coeffs: np.ndarray

    @property
    def order(self) -> int:
        return len(self.coeffs) - 1

    def _norm_coeffs(self) -> np.ndarray:
        k = np.arange(self.order + 1)
        return np.asarray(self.coeffs, dtype=np.float64) / _fact(k)

    def generate(self, t_arr: np.ndarray, t_ref: float = 0.0) -> np.ndarray:
        k = np.arange(self.order + 1)
        powers = (t_arr - t_ref)[None, :] ** k[:, None]
        delay = np.sum(self._norm_coeffs()[:, None] * powers, axis=0)
        return t_arr - delay / C_VAL

    def to_device(self, t_ref: float) -> PylModulation:
        nc = self._norm_coeffs()
        if len(nc) > PYL_MOD_MAX_COEFFS:
            msg = f"the GPU simulation takes at most {PYL_MOD_MAX_COEFFS} coefficients"
            raise ValueError(msg)
        return _packed(PYL_MOD_DERIVATIVE_SERIES, t_ref, coeffs=nc)


@dataclass(kw_only=True)
class CircularModulating:
    """Circular orbit (``modulate.py:122-150``): delay = x sin(omega (t - t_ref) + psi)
    light-seconds; x from the masses when not given."""

    p_orb: float
    psi: float
    x_orb: float | None = None
    m_c: float | None = None
    m_p: float = 1.4
    sin_i: float = 1.0

    def __post_init__(self) -> None:
        if self.x_orb is None and self.m_c is not None:
            a = 0.005 * ((self.m_p + self.m_c) * self.p_orb**2) ** (1 / 3)
            self.x_orb = a * (self.m_c / (self.m_p + self.m_c)) * self.sin_i

    def generate(self, t_arr: np.ndarray, t_ref: float = 0.0) -> np.ndarray:
        omega = 2 * np.pi / self.p_orb
        delay = self.x_orb * np.sin(omega * (t_arr - t_ref) + self.psi)
        return t_arr - delay

    def to_device(self, t_ref: float) -> PylModulation:
        return _packed(PYL_MOD_CIRCULAR, t_ref, (self.p_orb, self.psi, self.x_orb))

    def to_derivatives(self) -> dict[str, float]:
        """``modulate.py:152-171``: displacement derivatives d0..d5 at t_ref."""
        omega = 2.0 * np.pi / self.p_orb
        xs = self.x_orb * np.sin(self.psi)
        xc = self.x_orb * np.cos(self.psi)
        d2 = -C_VAL * xs * omega**2
        d3 = -C_VAL * xc * omega**3
        return {"shift": C_VAL * xs, "vel": C_VAL * xc * omega, "acc": d2, "jerk": d3,
                "snap": -d2 * omega**2, "crackle": -d3 * omega**2}


@dataclass(kw_only=True)
class CircularT0Modulating:
    """Circular orbit with an epoch (``modulate.py:214-226``); ``a`` in light-metres."""

    a: float
    p_orb: float
    t0: float

    def generate(self, t_arr: np.ndarray, t_ref: float = 0.0) -> np.ndarray:
        omega = 2 * np.pi / self.p_orb
        phi = 2 * np.pi * ((t_ref - self.t0) % self.p_orb) / self.p_orb
        delay = self.a * np.sin(omega * (t_arr - t_ref) + phi)
        return t_arr - delay / C_VAL

    def to_device(self, t_ref: float) -> PylModulation:
        return _packed(PYL_MOD_CIRCULAR_T0, t_ref, (self.a, self.p_orb, self.t0))


@dataclass(kw_only=True)
class KeplerianModulating:
    """Keplerian orbit (``modulate.py:229-250``, ``kepler.py:64-140``): Kepler's equation
    by 8 Newton steps from E = M, projected z in light-metres."""

    p_orb: float
    ecc: float
    phi: float
    a: float
    om: float
    inc: float

    def generate(self, t_arr: np.ndarray, t_ref: float = 0.0) -> np.ndarray:
        x = t_arr - t_ref
        ecc = abs(self.ecc)
        m_arr = 2 * np.pi * ((x % self.p_orb) / self.p_orb) + self.phi
        e_arr = m_arr.copy()
        for _ in range(8):
            e_arr += (m_arr + ecc * np.sin(e_arr) - e_arr) / (1 - ecc * np.cos(e_arr))
        nu = 2 * np.arctan(np.sqrt((1 + ecc) / (1 - ecc)) * np.tan(e_arr / 2))
        r = self.a * (1 - ecc**2) / (1 + ecc * np.cos(nu))
        return t_arr - r * np.sin(nu + self.om) * np.sin(self.inc) / C_VAL

    def to_device(self, t_ref: float) -> PylModulation:
        if not abs(self.ecc) < 1.0:
            msg = f"eccentricity must lie in [0, 1), got {self.ecc}"
            raise ValueError(msg)
        return _packed(PYL_MOD_KEPLERIAN, t_ref,
                       (self.p_orb, self.ecc, self.phi, self.a, self.om, self.inc))


type_to_mods: dict[str, type] = {
    "derivative": DerivativeModulating,
    "derivative_series": DerivativeSeriesModulating,
    "circular": CircularModulating,
    "circular_t0": CircularT0Modulating,
    "keplerian": KeplerianModulating,
}

__all__ = ["CircularModulating", "CircularT0Modulating", "DerivativeModulating",
           "DerivativeSeriesModulating", "KeplerianModulating", "PylModulation",
           "type_to_mods"]
