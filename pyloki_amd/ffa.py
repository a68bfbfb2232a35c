"""FFA driver: leaf brute fold + dynamic-programming levels on the GPU.

Mirrors ``pyloki.ffa`` (``src/pyloki/ffa.py``): ``DynamicProgramming`` keeps the
reference's attributes and methods (``initialize``, ``execute``, ``get_fold_norm``,
``fold``, ``param_arr``, ...). The fold lives in HBM as a torch tensor
(``fold_device``); ``fold`` returns a host copy in the reference's shape
``(nsegments, *param_counts, 2, nbins)``.

``FFAPlan`` holds everything that does not depend on the data: the per-level grids
(uploaded once), the brute-fold plan and the resolve tables. Series of different lengths
whose configurations share ``bseg_brute`` and the grids (as the default planning gives for
power-of-two lengths) share one plan and can be folded together by concatenating their
segments (``ffa_batch``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from pyloki_amd import kernels
from pyloki_amd.config import PulsarSearchConfig
from pyloki_amd.device import require_gpu, to_device
from pyloki_amd.io.timeseries import TimeSeries


@dataclass
class LevelPlan:
    level: int
    counts_prev: np.ndarray
    counts_cur: np.ndarray
    param_arr: list[np.ndarray]
    params_dev: torch.Tensor
    table: torch.Tensor
    ncoords: int
    phases: torch.Tensor | None = None  # FP64 resolved phases (complex folds shift by them)


class FFAPlan:
    """Data-independent plan of an FFA search (grids, resolve tables, brute-fold plan)."""

    def __init__(self, cfg: PulsarSearchConfig, nlevels: int | None = None) -> None:
        require_gpu()
        self.complex = bool(cfg.use_fourier)
        self.cfg = cfg
        self.nlevels = cfg.niters_ffa if nlevels is None else int(nlevels)
        dparams0 = cfg.get_dparams(0)
        self.param_arr0 = cfg.get_param_arr(dparams0)
        if cfg.nparams > 1:
            for i in range(cfg.nparams - 1):
                if len(self.param_arr0[i]) > 1:
                    msg = f"param_arr has {len(self.param_arr0[i])} values, should have only one"
                    raise ValueError(msg)
        self.counts0 = cfg.get_param_grid_count(0)
        self.t_ref0 = 0.0 if cfg.nparams == 1 else cfg.bseg_brute * cfg.tsamp / 2
        if self.complex:
            self.bf = kernels.ComplexFoldPlan(self.param_arr0[-1], cfg.bseg_brute, cfg.nbins,
                                              cfg.tsamp, self.t_ref0)
        else:
            self.bf = kernels.BruteFoldPlan(self.param_arr0[-1], cfg.bseg_brute, cfg.nbins,
                                            cfg.tsamp, self.t_ref0)
        self.levels: list[LevelPlan] = []
        counts_prev = self.counts0
        for level in range(1, self.nlevels + 1):
            pa = cfg.get_param_arr(cfg.get_dparams(level))
            counts_cur = np.array([len(a) for a in pa], dtype=np.int64)
            ncur = int(np.prod(counts_cur))
            params_dev = to_device(np.concatenate(pa).astype(np.float64))
            phases = (torch.empty((ncur, 2), dtype=torch.float64, device=params_dev.device)
                      if self.complex else None)
            table = kernels.ffa_resolve(params_dev, counts_cur, counts_prev, cfg.param_limits,
                                        level, cfg.tseg_brute, cfg.nbins, phases=phases)
            self.levels.append(LevelPlan(level, counts_prev, counts_cur, pa, params_dev, table,
                                         ncur, phases))
            counts_prev = counts_cur

    def boundary_counts(self) -> dict:
        """FP64 decisions of the plan that a +-1 ulp change of their input would flip:
        brute-fold phase bins, and the grid indices and bin shifts of every level's resolve
        (the honest bound on disagreement with an evaluation that rounds differently)."""
        from pyloki_amd import _lib
        from pyloki_amd.device import ptr, stream_ptr

        cnt = torch.zeros(2, dtype=torch.int64, device=self.levels[0].params_dev.device) \
            if self.levels else None
        lib = _lib.load()
        lim = np.ascontiguousarray(self.cfg.param_limits, dtype=np.float64).reshape(-1)
        n_idx = n_shift = 0
        for lp in self.levels:
            _lib.check(lib.pyl_ffa_resolve_boundary_count(
                len(lp.counts_cur), ptr(lp.params_dev), _lib.i64_array(lp.counts_cur),
                _lib.i64_array(lp.counts_prev), _lib.f64_array(lim), lp.level,
                self.cfg.tseg_brute, self.cfg.nbins, ptr(cnt), stream_ptr()))
            n_idx += 2 * lp.ncoords * int(np.sum(np.asarray(lp.counts_prev) > 1))
            n_shift += 2 * lp.ncoords
        got = cnt.cpu().numpy() if cnt is not None else np.zeros(2, dtype=np.int64)
        bf = 0 if self.complex else self.bf.boundary_count
        nbf = 0 if self.complex else self.bf.nfreq * self.cfg.bseg_brute
        return {"phase_bins": int(bf), "phase_bins_of": int(nbf),
                "grid_indices": int(got[0]), "grid_indices_of": int(n_idx),
                "shifts": int(got[1]), "shifts_of": int(n_shift)}

    def level(self, fold: torch.Tensor, lp: LevelPlan) -> torch.Tensor:
        """One level's shift-add (real, or complex with fractional phase shifts)."""
        if self.complex:
            return kernels.ffa_shift_add_complex(fold, lp.table, lp.phases, lp.ncoords,
                                                 self.cfg.nbins)
        return kernels.ffa_shift_add(fold, lp.table, lp.ncoords)

    def final_param_arr(self) -> list[np.ndarray]:
        return self.levels[-1].param_arr if self.levels else self.param_arr0

    def level_nseg(self, nsamps: int, level: int) -> int:
        return -(-nsamps // self.cfg.bseg_brute) // (2**level)

    def run(self, ts_e: torch.Tensor, ts_v: torch.Tensor, nlevels: int | None = None,
            keep: bool = False, dtype: torch.dtype = torch.float32
            ) -> torch.Tensor | list[torch.Tensor]:
        """Brute fold + ``nlevels`` levels; returns the last fold (or all, ``keep``).

        ``dtype`` selects the fold storage: FP32, or BF16 with FP32 accumulation."""
        nl = self.nlevels if nlevels is None else nlevels
        if self.complex:
            fold = self.bf.fold(ts_e, ts_v)
        else:
            fold = self.bf.fold(ts_e, ts_v, dtype=dtype)
        folds = [fold] if keep else None
        for lp in self.levels[:nl]:
            if fold.shape[0] < 2:
                msg = "not enough segments for another FFA level"
                raise ValueError(msg)
            fold = self.level(fold, lp)
            if keep:
                folds.append(fold)
        return folds if keep else fold


class DynamicProgramming:
    """FFA dynamic programming on the GPU; API of ``ffa.py:196-435``."""

    def __init__(self, ts_data: TimeSeries, cfg: PulsarSearchConfig) -> None:
        self._ts_data = ts_data
        self._cfg = cfg
        self._data_type = np.complex64 if cfg.use_fourier else np.float32
        self._plan: FFAPlan | None = None
        self._fold_dev: torch.Tensor | None = None
        self._fold_host: np.ndarray | None = None

    @property
    def ts_data(self) -> TimeSeries:
        return self._ts_data

    @property
    def cfg(self) -> PulsarSearchConfig:
        return self._cfg

    @property
    def data_type(self):
        return self._data_type

    @property
    def fold_device(self) -> torch.Tensor:
        """Current fold in HBM, flattened coords: ``[nseg, ncoords, 2, nbins]`` (complex:
        ``nbins // 2 + 1`` harmonics)."""
        return self._fold_dev

    @property
    def fold(self) -> np.ndarray:
        """Host copy of the fold in the reference's shape ``(nseg, *counts, 2, nbins)``."""
        if self._fold_host is None:
            f = self._fold_dev.cpu().numpy()
            counts = [len(p) for p in self._param_arr]
            self._fold_host = f.reshape(f.shape[0], *counts, *f.shape[-2:])
        return self._fold_host

    @property
    def nsegments(self) -> int:
        return self._fold_dev.shape[0]

    @property
    def nbins(self) -> int:
        return self._fold_dev.shape[-1]

    @property
    def ffa_level(self) -> int:
        return self._ffa_level

    @property
    def dparams(self) -> np.ndarray:
        return self._dparams

    @property
    def dparams_actual(self) -> np.ndarray:
        return self._dparams_actual

    @property
    def param_arr(self) -> list[np.ndarray]:
        return self._param_arr

    @property
    def param_grid_count(self) -> np.ndarray:
        return self._param_grid_count

    @property
    def param_arr_dict(self) -> dict[str, np.ndarray]:
        names = self.cfg.param_names
        return {f"{names[i]}": p for i, p in enumerate(self.param_arr)}

    @property
    def nparam_vol(self) -> int:
        return int(np.prod([len(p) for p in self.param_arr])) if self.param_arr else 0

    @property
    def leaves_lb(self) -> float:
        return np.round(np.log2(self.nparam_vol), 2)

    @property
    def tseg(self) -> float:
        return self._tseg

    def _set_fold(self, fold: torch.Tensor) -> None:
        self._fold_dev = fold
        self._fold_host = None

    def initialize(self) -> None:
        """Brute fold into level 0; ``ffa.py:305-329``."""
        self._ffa_level = 0
        self._plan = FFAPlan(self.cfg)
        ts_e = to_device(self.ts_data.ts_e, torch.float32)
        ts_v = to_device(self.ts_data.ts_v, torch.float32)
        self._set_fold(self._plan.bf.fold(ts_e, ts_v))
        self._param_arr = self._plan.param_arr0
        self._param_grid_count = self.cfg.get_param_grid_count(0)
        self._dparams = self.cfg.get_dparams(0)
        self._dparams_actual = self.cfg.get_dparams_actual(0)
        self._tseg = self.cfg.tseg_brute

    def execute(self, *, show_progress: bool = False) -> None:  # noqa: ARG002
        """Run every FFA level; ``ffa.py:331-347``."""
        for _ in range(self.cfg.niters_ffa):
            self._execute_iter()

    def _execute_iter(self) -> None:
        """One level; ``ffa.py:349-380``."""
        lp = self._plan.levels[self._ffa_level]
        self._ffa_level += 1
        self._set_fold(self._plan.level(self._fold_dev, lp))
        self._param_arr = lp.param_arr
        self._param_grid_count = self.cfg.get_param_grid_count(self._ffa_level)
        self._dparams = self.cfg.get_dparams(self._ffa_level)
        self._dparams_actual = self.cfg.get_dparams_actual(self._ffa_level)
        self._tseg *= 2

    def get_fold_norm(self, iseg: int = 0, param_idx: tuple | None = None) -> np.ndarray:
        """``e / sqrt(v)`` of one segment (host); ``ffa.py:382-411`` (complex folds are
        inverse-FFTed first)."""
        fold = self.fold
        if self.cfg.use_fourier:
            fold = np.fft.irfft(fold).astype(np.float32)
        sel = fold[iseg] if param_idx is None else fold[(iseg, *param_idx)]
        return sel[..., 0, :] / np.sqrt(sel[..., 1, :])

    def get_complexity(self) -> list[int]:
        out = []
        for level in range(self.cfg.niters_ffa + 1):
            dparams = self.cfg.get_dparams(level)
            out.append(int(np.prod([len(p) for p in self.cfg.get_param_arr(dparams)])))
        return out


def compute_ffa(tseries: TimeSeries, search_cfg: PulsarSearchConfig, *, quiet: bool = False,
                show_progress: bool = False) -> np.ndarray:  # noqa: ARG001
    """``ffa.py:438-467``."""
    dyp = DynamicProgramming(tseries, search_cfg)
    dyp.initialize()
    dyp.execute(show_progress=show_progress)
    return dyp.fold


def ffa_batch(plan: FFAPlan, series: list[tuple[torch.Tensor, torch.Tensor]],
              widths: np.ndarray, stdnoise: float = 1.0, *, fused: bool = True,
              dtype: torch.dtype = torch.float32) -> list[torch.Tensor]:
    """Fold and score many series of (possibly) different lengths in one pass per level.

    All series share ``plan`` (same ``bseg_brute`` and grids); series are ordered longest
    first and concatenated along the segment axis, so level L works on every series that
    still has more than one segment in one launch. A series whose last level this is gets
    that level fused with its scoring (``pyl_ffa_shift_add_score``), so its final fold is
    never written; the level before it stores its output in the "wrap" row layout, which
    the fused kernel reads without rotation selects. With ``fused=False`` the final fold is
    stored and scored separately.
    Returns the per-width S/N grids ``[ncoords_final(series), n_w]`` in input order
    (``ffa_search``'s grid, ``search.py:46-47``).
    """
    if plan.complex:
        msg = "ffa_batch runs real folds; use DynamicProgramming for use_fourier=True"
        raise NotImplementedError(msg)
    nser = len(series)
    lens = [int(e.shape[0]) for e, _ in series]
    order = sorted(range(nser), key=lambda i: -lens[i])
    for i in range(nser):
        if lens[i] % plan.cfg.bseg_brute:
            msg = "series length must be a multiple of bseg_brute"
            raise ValueError(msg)
    nseg = [lens[i] // plan.cfg.bseg_brute for i in order]
    for n in nseg:
        if n & (n - 1):
            msg = "segments per series must be a power of two"
            raise ValueError(msg)
    if int(math.log2(max(nseg))) > plan.nlevels:
        msg = "plan has fewer levels than the longest series needs"
        raise ValueError(msg)
    dev = series[0][0].device
    fold = torch.empty((sum(nseg), plan.bf.nfreq, 2, plan.cfg.nbins), dtype=dtype, device=dev)
    plan.bf.fold_batch([series[i] for i in order], fold)
    results: dict[int, torch.Tensor] = {}
    level = 0
    nbins = plan.cfg.nbins
    use_fused = fused and nbins in (32, 64, 128)
    # The level before a series' last one writes the "wrap" row layout (FP32), which lets
    # the fused last level load every rotated 4-bin chunk with one contiguous load.
    pre_layout = "wrap" if use_fused and dtype == torch.float32 else "dense"
    # A series with a single segment from the start is scored at level 0.
    off = sum(n for n in nseg if n > 1)
    for k in range(nser):
        if nseg[k] == 1:
            results[order[k]] = kernels.score_boxcar(fold[off], widths, stdnoise)
            off += 1
    if use_fused and nbins in (32, 64) and len(widths) <= 16 and dtype == torch.float32:
        _batch_two_levels(plan, fold, nseg, order, widths, stdnoise, results)
        return [results[i] for i in range(nser)]
    fin_src: torch.Tensor | None = None  # output of the previous level's "pre" group
    while True:
        big = [k for k in range(nser) if nseg[k] > 4]
        pre = [k for k in range(nser) if nseg[k] == 4]
        fin = [k for k in range(nser) if nseg[k] == 2]
        if not big and not pre and not fin:
            break
        lp = plan.levels[level]
        nb = sum(nseg[k] for k in big)
        npre = 4 * len(pre)
        if fin:
            if fin_src is not None:
                part, lay = fin_src, pre_layout
            else:  # series that had 2 segments after the brute fold
                part, lay = fold[nb + npre:nb + npre + 2 * len(fin)], "dense"
            if use_fused:
                snr = kernels.ffa_shift_add_score(part, lp.table, lp.ncoords, widths, stdnoise,
                                                  n_inner=int(lp.counts_cur[-1]),
                                                  in_layout=lay)
                for m, k in enumerate(fin):
                    results[order[k]] = snr[m]
            else:
                last = kernels.ffa_shift_add(part, lp.table, lp.ncoords, in_layout=lay)
                for m, k in enumerate(fin):
                    results[order[k]] = kernels.score_boxcar(last[m], widths, stdnoise)
        fin_src = None
        if pre:
            fin_src = kernels.ffa_shift_add(fold[nb:nb + npre], lp.table, lp.ncoords,
                                            out_layout=pre_layout)
        if big:
            fold = kernels.ffa_shift_add(fold[:nb], lp.table, lp.ncoords)
        for k in big + pre + fin:
            nseg[k] //= 2
        level += 1
    return [results[i] for i in range(nser)]


def _batch_two_levels(plan: FFAPlan, fold: torch.Tensor, nseg: list[int], order: list[int],
                      widths: np.ndarray, stdnoise: float, results: dict) -> None:
    """The level loop of ``ffa_batch`` for FP32 folds with nbins 32/64 and at most 16
    widths. Series of one length (contiguous in ``fold``, longest first) run together:
    their levels two at a time (``pyl_ffa_shift_add2_layout``, the intermediate fold never
    stored; an odd count ends with one single level), every fold after the brute fold in
    the "wrap" row layout (each rotated 4-bin chunk one load), and the last two levels
    fused with scoring (``pyl_ffa_shift_add2_score_layout``). A series with 2
    segments after the brute fold uses the one-level fused kernel; the series with one
    segment have been scored already."""
    nser = len(nseg)
    off, k = 0, 0
    while k < nser:
        n0 = nseg[k]
        k1 = k
        while k1 < nser and nseg[k1] == n0:
            k1 += 1
        part = fold[off:off + (k1 - k) * n0]
        off += (k1 - k) * n0
        if n0 > 1:
            snr = _class_levels(plan, part, n0, widths, stdnoise)
            for m in range(k1 - k):
                results[order[k + m]] = snr[m]
        k = k1


def _class_levels(plan: FFAPlan, fold: torch.Tensor, n0: int, widths: np.ndarray,
                  stdnoise: float) -> torch.Tensor:
    """Every level of series with ``n0`` segments each (concatenated in ``fold``) and the
    per-width S/N of their last level: [nseries, ncoords_final, n_w]."""
    left = int(math.log2(n0))
    level, lay = 0, "dense"
    while left > 2:  # every level written in the wrap layout, read with one load per chunk
        if left == 3:
            lp = plan.levels[level]
            fold = kernels.ffa_shift_add(fold, lp.table, lp.ncoords, in_layout=lay,
                                         out_layout="wrap")
            level, left = level + 1, left - 1
        else:
            a, b = plan.levels[level], plan.levels[level + 1]
            fold = kernels.ffa_shift_add2(fold, a.table, b.table, a.ncoords, b.ncoords,
                                          in_layout=lay, out_layout="wrap")
            level, left = level + 2, left - 2
        lay = "wrap"
    if left == 2:
        a, b = plan.levels[level], plan.levels[level + 1]
        return kernels.ffa_shift_add2_score(fold, a.table, b.table, a.ncoords, b.ncoords,
                                            widths, stdnoise, n_inner=int(b.counts_cur[-1]),
                                            in_layout=lay)
    lp = plan.levels[level]
    return kernels.ffa_shift_add_score(fold, lp.table, lp.ncoords, widths, stdnoise,
                                       n_inner=int(lp.counts_cur[-1]), in_layout=lay)
