"""Search configuration and per-level grid planning.

Mirrors ``pyloki.config`` (``src/pyloki/config.py``): the same class names, fields,
validators, defaults and grid formulas, so that notebooks written against pyloki build
the identical plan here. Everything in this module is host-side FP64 planning; the
per-leaf work it plans runs in the HIP library.
"""

from __future__ import annotations

import itertools
import math
from dataclasses import dataclass, field

import numpy as np

from pyloki_amd.detection.scoring import generate_box_width_trials
from pyloki_amd.utils import psr_utils, transforms
from pyloki_amd.utils.misc import C_VAL, G_CONST, M_SUN


def _is_power_of_two(n: int) -> bool:
    return n > 0 and (n & (n - 1)) == 0


class ParamLimits:
    """Search parameter limits, ordered ``..., jerk, accel, freq``.

    Restates ``ParamLimits`` (``config.py:59-321``) for the builders the notebooks use.
    """

    def __init__(self, limits: np.ndarray | list[tuple[float, float]]) -> None:
        self.limits = np.asarray(limits, dtype=np.float64)

    def get_cheby_limits(self, tobs: float) -> list[tuple[float, float]]:
        """Chebyshev coefficient bounds from the Taylor corners; ``config.py:72-104``."""
        t_s = tobs / 2
        d_limits = self.limits[:-1][::-1]
        d_limits = [(0, 0), (0, 0), *d_limits]
        d_corners = np.array(list(itertools.product(*list(d_limits))))
        alpha_corners = np.vstack(
            [transforms.taylor_to_cheby(d_vec, t_s) for d_vec in d_corners],
        )
        alpha_bounds = list(
            zip(np.min(alpha_corners, axis=0), np.max(alpha_corners, axis=0)),
        )
        return [tuple(self.limits[-1]), *alpha_bounds]

    @classmethod
    def from_taylor(
        cls,
        freq: tuple[float, float],
        accel: tuple[float, float] | None = None,
        jerk: tuple[float, float] | None = None,
        snap: tuple[float, float] | None = None,
    ) -> ParamLimits:
        """``config.py:106-144``."""
        default_limit = (0.0, 0.0)
        all_params = [snap, jerk, accel, freq]
        last_non_none = next(i for i, p in enumerate(all_params) if p is not None)
        out = [p if p is not None else default_limit for p in all_params[last_non_none:]]
        return cls([(float(a), float(b)) for a, b in out])

    @classmethod
    def from_upper(
        cls,
        freq: float | tuple[float, float],
        true_d_params: list[float],
        d_upper_limits: tuple[float, float],
        t_obs: float,
    ) -> ParamLimits:
        """Limits from the range of the highest derivative; ``config.py:146-197``."""
        nparams = len(true_d_params) + 1
        dvec = np.zeros(nparams + 1, dtype=np.float64)
        dvec[1:-2] = true_d_params[1:]
        dvec[0] = d_upper_limits[0]
        dvec_min_up = transforms.shift_taylor_params(dvec, t_obs / 2)
        dvec_min_low = transforms.shift_taylor_params(dvec, -t_obs / 2)
        dvec[0] = d_upper_limits[1]
        dvec_max_up = transforms.shift_taylor_params(dvec, t_obs / 2)
        dvec_max_low = transforms.shift_taylor_params(dvec, -t_obs / 2)
        dvec_bound_low = np.minimum(dvec_min_low, dvec_max_low)
        dvec_bound_up = np.maximum(dvec_min_up, dvec_max_up)
        bounds_d = list(zip(dvec_bound_low, dvec_bound_up))
        bounds = [(float(a), float(b)) for a, b in bounds_d[:-2]]
        freq_shift = dvec_bound_up[-2] / C_VAL
        if isinstance(freq, (tuple, list, np.ndarray)):
            freq_min, freq_max = freq
            bounds.append((freq_min * (1 - freq_shift), freq_max * (1 + freq_shift)))
        else:
            bounds.append((freq * (1 - freq_shift), freq * (1 + freq_shift)))
        return cls(bounds)

    @classmethod
    def from_circular(
        cls,
        freq: float | tuple[float, float],
        p_orb_min: float,
        m_c_max: float = 10.0,
        m_p_min: float = 1.2,
        poly_order: int = 5,
    ) -> ParamLimits:
        """Limits from circular-orbit bounds; ``config.py:199-251``."""
        poly_order = max(poly_order, 2)
        omega_orb_max = 2 * np.pi / p_orb_min
        x_const = float(
            (M_SUN * G_CONST) ** (1 / 3) / (C_VAL * (2 * np.pi) ** (2 / 3)),
        )
        x_orb = (
            x_const
            * ((m_p_min + m_c_max) * p_orb_min**2) ** (1 / 3)
            * m_c_max
            / (m_p_min + m_c_max)
        )
        max_derivs = x_orb * C_VAL * omega_orb_max ** np.arange(poly_order + 1)
        bounds = [(-d, d) for d in max_derivs[2:][::-1]]
        freq_shift = max_derivs[1] / C_VAL
        if isinstance(freq, (tuple, list, np.ndarray)):
            freq_min, freq_max = freq
            bounds.append((freq_min * (1 - freq_shift), freq_max * (1 + freq_shift)))
        else:
            bounds.append((freq * (1 - freq_shift), freq * (1 + freq_shift)))
        return cls(bounds)


_TILING = ("aggressive", "quadrature", "conservative")


@dataclass(frozen=True, kw_only=True)
class PulsarSearchConfig:
    """Configuration of a polynomial (Taylor) search; ``config.py:324-436``.

    Parameter limits are ordered ``..., jerk, accel, freq``. ``bseg_brute`` and
    ``bseg_ffa`` default (when 0) to the reference's formulas
    (``_bseg_brute_default``/``_bseg_ffa_default``, ``config.py:821-827``).
    """

    nsamps: int
    tsamp: float
    nbins: int
    eta: float
    param_limits: np.ndarray
    ducy_max: float = 0.2
    wtsp: float = 1.5
    prune_poly_order: int = 3
    p_orb_min: float = 0.0
    m_c_max: float = 10.0
    m_p_min: float = 1.4
    bseg_brute: int = 0
    bseg_ffa: int = 0
    use_fourier: bool = True
    branch_max: int = 16
    propagator_significance: float = 2.0
    validation_significance: float = 5.0
    tiling_strategy: str = "aggressive"
    _cache: dict = field(default_factory=dict, init=False, repr=False, compare=False)

    def __post_init__(self) -> None:
        lim = np.asarray(self.param_limits, dtype=np.float64)
        if lim.ndim != 2 or lim.shape[1] != 2 or lim.shape[0] < 1:
            msg = f"param_limits must have 1-5 elements, got {lim.shape[0] if lim.ndim else 0}"
            raise ValueError(msg)
        for i in range(lim.shape[0]):
            if lim[i, 0] >= lim[i, 1]:
                msg = f"param_limits must have min < max, got {lim}"
                raise ValueError(msg)
        object.__setattr__(self, "param_limits", lim)
        if not isinstance(self.nsamps, (int, np.integer)) or self.nsamps <= 0:
            msg = f"'nsamps' must be a positive int: {self.nsamps}"
            raise ValueError(msg)
        if not _is_power_of_two(int(self.nsamps)):
            msg = f"'nsamps' must be a power of 2: {self.nsamps}"
            raise ValueError(msg)
        if not self.tsamp > 0:
            msg = f"'tsamp' must be > 0: {self.tsamp}"
            raise ValueError(msg)
        if not isinstance(self.nbins, (int, np.integer)) or self.nbins <= 0:
            msg = f"'nbins' must be a positive int: {self.nbins}"
            raise ValueError(msg)
        if not self.eta > 0:
            msg = f"'eta' must be > 0: {self.eta}"
            raise ValueError(msg)
        if not self.ducy_max > 0 or not self.wtsp > 0:
            msg = "'ducy_max' and 'wtsp' must be > 0"
            raise ValueError(msg)
        if self.prune_poly_order <= 1:
            msg = f"'prune_poly_order' must be > 1: {self.prune_poly_order}"
            raise ValueError(msg)
        for name in ("bseg_brute", "bseg_ffa"):
            val = int(getattr(self, name))
            if val != 0 and not _is_power_of_two(val):
                msg = f"'{name}' must be a power of 2: {val}"
                raise ValueError(msg)
        if self.branch_max <= 10:
            msg = f"'branch_max' must be > 10: {self.branch_max}"
            raise ValueError(msg)
        if self.tiling_strategy not in _TILING:
            msg = f"'tiling_strategy' must be in {_TILING}: {self.tiling_strategy}"
            raise ValueError(msg)
        if self.bseg_brute == 0:
            object.__setattr__(self, "bseg_brute", self._bseg_brute_default())
        if self.bseg_ffa == 0:
            object.__setattr__(self, "bseg_ffa", self._bseg_ffa_default())
        if self.bseg_brute > self.nsamps:
            msg = f"bseg_brute ({self.bseg_brute}) must be < nsamps ({self.nsamps})"
            raise ValueError(msg)
        if self.bseg_ffa > self.nsamps:
            msg = f"bseg_ffa ({self.bseg_ffa}) must be <= nsamps ({self.nsamps})"
            raise ValueError(msg)
        if self.bseg_ffa <= self.bseg_brute:
            msg = f"bseg_ffa ({self.bseg_ffa}) must be > bseg_brute ({self.bseg_brute})"
            raise ValueError(msg)
        if self.prune_poly_order == 5 and self.p_orb_min == 0:
            msg = "p_orb_min must be provided for a circular orbit search"
            raise ValueError(msg)

    @property
    def tseg_brute(self) -> float:
        return self.bseg_brute * self.tsamp

    @property
    def tseg_ffa(self) -> float:
        return self.bseg_ffa * self.tsamp

    @property
    def niters_ffa(self) -> int:
        return int(np.log2(self.bseg_ffa / self.bseg_brute))

    @property
    def nparams(self) -> int:
        return self.param_limits.shape[0]

    @property
    def param_names(self) -> list[str]:
        return ["crackle", "snap", "jerk", "accel", "freq"][-self.nparams :]

    @property
    def f_min(self) -> float:
        return float(self.param_limits[-1][0])

    @property
    def f_max(self) -> float:
        return float(self.param_limits[-1][1])

    @property
    def x_mass_const(self) -> float:
        """``config.py:474-481`` (inflated by 10% for safety)."""
        k_const = float((M_SUN * G_CONST) ** (1 / 3))
        return k_const * 1.1 * self.m_c_max / (self.m_p_min + self.m_c_max) ** (2 / 3)

    @property
    def score_widths(self) -> np.ndarray:
        return generate_box_width_trials(self.nbins, ducy_max=self.ducy_max, wtsp=self.wtsp)

    def get_dparams_f(self, ffa_level: int) -> np.ndarray:
        tseg_cur = 2**ffa_level * self.tseg_brute
        t_ref = 0 if self.nparams == 1 else tseg_cur / 2
        return psr_utils.poly_taylor_step_f(self.nparams, tseg_cur, self.nbins, self.eta, t_ref)

    def get_dparams(self, ffa_level: int, use_cheby_coarsening: bool = True) -> np.ndarray:
        """Step sizes at an FFA level; ``config.py:515-540``."""
        tseg_cur = 2**ffa_level * self.tseg_brute
        t_ref = 0 if self.nparams == 1 else tseg_cur / 2
        return psr_utils.poly_taylor_step_d_f(
            self.nparams,
            tseg_cur,
            self.nbins,
            self.eta,
            self.f_max,
            t_ref=t_ref,
            use_cheby=use_cheby_coarsening,
        )

    def get_dparams_actual(self, ffa_level: int, use_cheby_coarsening: bool = True) -> np.ndarray:
        """Physically tiled spacings ``(vmax - vmin) / count``; ``config.py:542-570``."""
        counts = self.get_param_grid_count(ffa_level, use_cheby_coarsening)
        out = np.zeros(self.nparams, dtype=np.float64)
        for i in range(self.nparams):
            vmin, vmax = self.param_limits[i]
            out[i] = (vmax - vmin) / counts[i]
        return out

    def get_param_grid_count(self, ffa_level: int, use_cheby_coarsening: bool = True) -> np.ndarray:
        """Grid counts at an FFA level; ``config.py:572-597``."""
        dparams = self.get_dparams(ffa_level, use_cheby_coarsening)
        count = np.empty(self.nparams, dtype=np.int64)
        for i in range(self.nparams):
            count[i] = psr_utils.range_param_count(
                float(self.param_limits[i, 0]),
                float(self.param_limits[i, 1]),
                float(dparams[i]),
            )
        return count

    def get_param_arr(self, dparams: np.ndarray) -> list[np.ndarray]:
        """Grid cell centres for given steps; ``config.py:599-634``."""
        if len(dparams) != self.nparams:
            msg = f"dparams must have length {self.nparams}, got {len(dparams)}"
            raise ValueError(msg)
        return [
            psr_utils.range_param(
                float(self.param_limits[i, 0]),
                float(self.param_limits[i, 1]),
                float(dparams[i]),
            )
            for i in range(self.nparams)
        ]

    def generate_branching_pattern(
        self,
        kind: str = "poly_taylor_moving",
        ref_seg: int = 0,
        use_cheby_coarsening: bool = True,
        *,
        device: bool = False,
    ) -> np.ndarray:
        """Expected children per parent at every pruning level (``config.py:758-819``).

        Taylor kinds restate ``generate_bp_poly_taylor`` (``core/taylor.py:414-499``),
        Chebyshev kinds ``generate_bp_poly_chebyshev`` (``core/chebyshev.py:458-546``) and
        circular kinds ``generate_bp_circ_taylor`` (``core/circular.py:905-1010``): per
        final-grid frequency, the children per parent of each level, averaged with weights
        equal to the running product of earlier branchings. ``device=True`` runs the
        per-frequency recurrence of the Taylor and circular kinds on the GPU
        (``pyl_bp_counts``; the same values bit for bit).
        """
        funcs = {"poly_taylor_moving": (generate_bp_poly_taylor, True),
                 "poly_taylor_fixed": (generate_bp_poly_taylor, False),
                 "poly_chebyshev_moving": (generate_bp_poly_chebyshev, True),
                 "poly_chebyshev_fixed": (generate_bp_poly_chebyshev, False),
                 "circ_taylor_moving": (generate_bp_circ_taylor, True),
                 "circ_taylor_fixed": (generate_bp_circ_taylor, False)}
        if kind not in funcs:
            msg = f"Invalid kind: {kind}"
            raise ValueError(msg)
        if kind.startswith("poly_chebyshev"):
            use_cheby_coarsening = True
        func, use_moving_grid = funcs[kind]
        nsegments = int(np.ceil(self.nsamps / self.bseg_ffa))
        dparams = self.get_dparams(self.niters_ffa, use_cheby_coarsening)
        param_arr = self.get_param_arr(dparams)
        dparams_act = self.get_dparams_actual(self.niters_ffa, use_cheby_coarsening)
        if device:
            if kind.startswith("poly_chebyshev"):
                msg = "the device pattern covers the Taylor and circular kinds"
                raise ValueError(msg)
            return generate_bp_device(
                param_arr, dparams_act, self.tseg_ffa, nsegments, self.nbins, self.eta,
                ref_seg, use_moving_grid=use_moving_grid, tiling_strategy=self.tiling_strategy,
                use_cheby_coarsening=use_cheby_coarsening,
                circular=kind.startswith("circ"))
        return func(
            param_arr, dparams_act, self.tseg_ffa, nsegments, self.nbins, self.eta, ref_seg,
            use_moving_grid=use_moving_grid, tiling_strategy=self.tiling_strategy,
            use_cheby_coarsening=use_cheby_coarsening,
        )

    def _bseg_brute_default(self) -> int:
        init_levels = 1 if self.nparams == 1 else 5
        levels = int(np.log2(self.nsamps * self.tsamp * self.f_min) - init_levels)
        return int(self.nsamps / 2**levels)

    def _bseg_ffa_default(self) -> int:
        return int(self.nsamps)


def generate_bp_poly_taylor(
    param_arr: list[np.ndarray],
    dparams_act: np.ndarray,
    tseg_ffa: float,
    nsegments: int,
    nbins: int,
    eta: float,
    ref_seg: int,
    *,
    use_moving_grid: bool,
    tiling_strategy: str,
    use_cheby_coarsening: bool = True,
) -> np.ndarray:
    """Exact Taylor branching pattern (``core/taylor.py:414-499``), host FP64."""
    from pyloki_amd.utils.misc import FLOAT_EPSILON
    from pyloki_amd.utils.snail import MiddleOutScheme

    n_params = len(dparams_act)
    f0 = np.asarray(param_arr[-1], dtype=np.float64)
    nf = len(f0)
    scheme = MiddleOutScheme(nsegments, ref_seg, tseg_ffa)
    weights = np.ones(nf)
    bp = np.empty(nsegments - 1)
    dcur = np.tile(np.asarray(dparams_act, dtype=np.float64), (nf, 1))
    dcur[:, n_params - 1] *= C_VAL / f0
    for level in range(1, nsegments):
        coord_next = scheme.get_coord(level)
        coord_cur = scheme.get_current_coord(level, use_moving_grid)
        dnew, shift = _taylor_widths_step(dcur, f0, coord_cur[1], nbins, eta,
                                          use_cheby_coarsening)
        nbr = np.ones(nf, dtype=np.int64)
        dnext = dcur.copy()
        split = shift >= (eta - FLOAT_EPSILON)
        for i in range(nf):
            for j in range(n_params):
                if not split[i, j]:
                    continue
                num = max(1, math.ceil(dcur[i, j] / dnew[i, j] - FLOAT_EPSILON))
                nbr[i] *= num
                dnext[i, j] = dcur[i, j] / num
        bp[level - 1] = np.sum(weights * nbr) / np.sum(weights)
        weights = weights * nbr
        if use_moving_grid:
            dt = coord_next[0] - coord_cur[0]
            dvec = np.zeros((nf, n_params + 1))
            dvec[:, :-1] = dnext
            dcur = transforms.shift_taylor_errors(dvec, dt, tiling_strategy)[:, :-1]
        else:
            dcur = dnext
    return bp


def _taylor_widths_step(dcur: np.ndarray, f0: np.ndarray, tobs: float, nbins: int,
                        eta: float, use_cheby_coarsening: bool
                        ) -> tuple[np.ndarray, np.ndarray]:
    """Target widths of ``[d_k .. d_1]`` at span ``tobs`` (``psr_utils.py:132-143``) and the
    bin shift the current widths ``dcur`` make there (``psr_utils.py:228-248``)."""
    n_params = dcur.shape[1]
    dpf = psr_utils.poly_taylor_step_f(n_params, tobs, nbins, eta, 0.0, use_cheby_coarsening)
    dnew = dpf[None, :] * C_VAL / f0[:, None]
    k = np.arange(n_params - 1, -1, -1)
    fac = np.array([psr_utils.powi(tobs, int(kk) + 1) * nbins / psr_utils.fact(int(kk) + 1)
                    for kk in k])
    if use_cheby_coarsening:
        fac = fac / (2.0**k)
    return dnew, np.abs(dcur - dnew) * (fac[None, :] * (f0 / C_VAL)[:, None])


def generate_bp_circ_taylor(
    param_arr: list[np.ndarray],
    dparams_act: np.ndarray,
    tseg_ffa: float,
    nsegments: int,
    nbins: int,
    eta: float,
    ref_seg: int,
    *,
    use_moving_grid: bool,
    tiling_strategy: str,
    use_cheby_coarsening: bool = True,
) -> np.ndarray:
    """Circular-orbit Taylor branching pattern (``core/circular.py:905-1010``), host FP64.

    As the Taylor pattern over ``[crackle, snap, jerk, accel, v]``, except that: the
    crackle width never splits (the circular branch derives crackle from the other
    terms); branch counts stay floating point; and at the first level where a frequency's
    snap splits, its count is halved, the expected share of children whose snap/accel
    pair validates as an orbit at the level where it is first resolved. Five parameters
    are required. The reference leaves the crackle entry of the split widths
    uninitialised; it is never read (the loop over splits starts at snap, and the
    default "aggressive" propagation is diagonal), so it keeps its current width here.
    """
    from pyloki_amd.utils.misc import FLOAT_EPSILON
    from pyloki_amd.utils.snail import MiddleOutScheme

    n_params = len(dparams_act)
    if n_params != 5:
        msg = "Circular branching pattern requires exactly 5 parameters."
        raise ValueError(msg)
    f0 = np.asarray(param_arr[-1], dtype=np.float64)
    nf = len(f0)
    scheme = MiddleOutScheme(nsegments, ref_seg, tseg_ffa)
    weights = np.ones(nf)
    bp = np.empty(nsegments - 1)
    dcur = np.tile(np.asarray(dparams_act, dtype=np.float64), (nf, 1))
    dcur[:, n_params - 1] *= C_VAL / f0
    snap_seen = np.zeros(nf, dtype=bool)
    for level in range(1, nsegments):
        coord_next = scheme.get_coord(level)
        coord_cur = scheme.get_current_coord(level, use_moving_grid)
        dnew, shift = _taylor_widths_step(dcur, f0, coord_cur[1], nbins, eta,
                                          use_cheby_coarsening)
        nbr = np.ones(nf)
        snap_split = np.zeros(nf, dtype=bool)
        dnext = dcur.copy()
        for i in range(nf):
            for j in range(1, n_params):
                if shift[i, j] < (eta - FLOAT_EPSILON):
                    continue
                num = max(1, math.ceil(dcur[i, j] / dnew[i, j] - FLOAT_EPSILON))
                nbr[i] *= num
                dnext[i, j] = dcur[i, j] / num
                if j == 1 and num > 1:
                    snap_split[i] = True
        first = snap_split & ~snap_seen
        nbr[first] *= 0.5
        snap_seen |= first
        bp[level - 1] = np.sum(weights * nbr) / np.sum(weights)
        weights = weights * nbr
        if use_moving_grid:
            dvec = np.zeros((nf, n_params + 1))
            dvec[:, :-1] = dnext
            dcur = transforms.shift_taylor_errors(dvec, coord_next[0] - coord_cur[0],
                                                  tiling_strategy)[:, :-1]
        else:
            dcur = dnext
    return bp


def generate_bp_device(param_arr: list[np.ndarray], dparams_act: np.ndarray, tseg_ffa: float,
                       nsegments: int, nbins: int, eta: float, ref_seg: int, *,
                       use_moving_grid: bool, tiling_strategy: str,
                       use_cheby_coarsening: bool = True, circular: bool = False
                       ) -> np.ndarray:
    """``generate_bp_poly_taylor`` / ``generate_bp_circ_taylor`` with the per-frequency
    width recurrence on the GPU (``pyl_bp_counts``): the host supplies each level's
    constants (the step, the shift factor and the Taylor shift matrix of the moving grid,
    from the same helpers as the host generators) and forms the weighted averages with
    the same sums, so the pattern equals the host one bit for bit."""
    import ctypes

    from pyloki_amd import _lib
    from pyloki_amd.device import require_gpu, stream_ptr
    from pyloki_amd.utils.snail import MiddleOutScheme

    require_gpu()
    n_params = len(dparams_act)
    f0 = np.ascontiguousarray(param_arr[-1], dtype=np.float64)
    nf = len(f0)
    nlev = nsegments - 1
    scheme = MiddleOutScheme(nsegments, ref_seg, tseg_ffa)
    dpf = np.empty((max(nlev, 1), n_params))
    fac = np.empty((max(nlev, 1), n_params))
    tm = np.zeros((max(nlev, 1), n_params + 1, n_params + 1))
    k = np.arange(n_params - 1, -1, -1)
    for level in range(1, nsegments):
        coord_next = scheme.get_coord(level)
        coord_cur = scheme.get_current_coord(level, use_moving_grid)
        tobs = coord_cur[1]
        dpf[level - 1] = psr_utils.poly_taylor_step_f(n_params, tobs, nbins, eta, 0.0,
                                                      use_cheby_coarsening)
        fc = np.array([psr_utils.powi(tobs, int(kk) + 1) * nbins / psr_utils.fact(int(kk) + 1)
                       for kk in k])
        if use_cheby_coarsening:
            fc = fc / (2.0**k)
        fac[level - 1] = fc
        if use_moving_grid:
            tm[level - 1] = transforms.taylor_shift_matrix(n_params + 1,
                                                           coord_next[0] - coord_cur[0])
    tiling = {"aggressive": 0, "conservative": 1, "quadrature": 2}.get(tiling_strategy)
    if tiling is None:
        msg = f"Invalid tiling strategy: {tiling_strategy}"
        raise ValueError(msg)
    nbr = np.empty((max(nlev, 1), nf))
    dinit = np.ascontiguousarray(dparams_act, dtype=np.float64)
    p64 = ctypes.POINTER(ctypes.c_double)
    _lib.check(_lib.load().pyl_bp_counts(
        1 if circular else 0, n_params, nlev, nf, f0.ctypes.data_as(p64),
        dinit.ctypes.data_as(p64), np.ascontiguousarray(dpf).ctypes.data_as(p64),
        np.ascontiguousarray(fac).ctypes.data_as(p64),
        np.ascontiguousarray(tm).ctypes.data_as(p64), int(use_moving_grid), tiling, float(eta),
        nbr.ctypes.data_as(p64), stream_ptr()))
    weights = np.ones(nf)
    bp = np.empty(nlev)
    for level in range(nlev):
        bp[level] = np.sum(weights * nbr[level]) / np.sum(weights)
        weights = weights * nbr[level]
    return bp


def ffa_level_counts(cfg: PulsarSearchConfig) -> list[np.ndarray]:
    """Grid counts for every FFA level 0..niters_ffa (planning helper)."""
    return [cfg.get_param_grid_count(level) for level in range(cfg.niters_ffa + 1)]


def ncoords(counts: np.ndarray) -> int:
    return int(math.prod(int(c) for c in counts))


def generate_bp_poly_chebyshev(
    param_arr: list[np.ndarray],
    dparams_act: np.ndarray,
    tseg_ffa: float,
    nsegments: int,
    nbins: int,
    eta: float,
    ref_seg: int,
    *,
    use_moving_grid: bool,
    tiling_strategy: str,
    use_cheby_coarsening: bool = True,  # noqa: ARG001
) -> np.ndarray:
    """Exact Chebyshev branching pattern (``core/chebyshev.py:458-546``), host FP64.

    The cell widths of every Chebyshev coefficient (alpha_0 included) are moved from the
    previous to the current domain, split where the step ``eta/nbins * C / f0`` is exceeded
    by a bin, then moved on to the next domain. The reference leaves the alpha_0 width of
    the split result uninitialised; it is never read and no other width depends on it, so
    it is kept at 0 here.
    """
    from pyloki_amd.utils.misc import FLOAT_EPSILON
    from pyloki_amd.utils.snail import MiddleOutScheme

    n_params = len(dparams_act)
    f0 = np.asarray(param_arr[-1], dtype=np.float64)
    nf = len(f0)
    scheme = MiddleOutScheme(nsegments, ref_seg, tseg_ffa)
    weights = np.ones(nf)
    bp = np.empty(nsegments - 1)
    dfull = np.zeros((nf, n_params + 1))
    dfull[:, :n_params] = np.asarray(dparams_act, dtype=np.float64)[None, :]
    dfull[:, n_params - 1] *= C_VAL / f0
    dcur = transforms.taylor_to_cheby_errors(dfull, scheme.get_coord(0)[1])
    dphi = eta / nbins
    for level in range(1, nsegments):
        coord_next = scheme.get_coord(level)
        coord_cur = scheme.get_current_coord(level, use_moving_grid)
        coord_prev = scheme.get_previous_coord(level, use_moving_grid)
        dcur = transforms.shift_cheby_errors(dcur, coord_cur, coord_prev, tiling_strategy)
        dnew = np.empty_like(dcur)
        for j in range(n_params + 1):
            dnew[:, j] = dphi * C_VAL / f0
        shift = np.abs(dcur - dnew) * (nbins * (f0 / C_VAL))[:, None]
        nbr = np.ones(nf, dtype=np.int64)
        dnext = np.zeros_like(dcur)
        for i in range(nf):
            for j in range(n_params):
                if shift[i, j] < (eta - FLOAT_EPSILON):
                    dnext[i, j] = dcur[i, j]
                    continue
                num = max(1, math.ceil(dcur[i, j] / dnew[i, j] - FLOAT_EPSILON))
                nbr[i] *= num
                dnext[i, j] = dcur[i, j] / num
        bp[level - 1] = np.sum(weights * nbr) / np.sum(weights)
        weights = weights * nbr
        if use_moving_grid:
            dnext = transforms.shift_cheby_errors(dnext, coord_next, coord_cur, tiling_strategy)
        dcur = dnext
    return bp
