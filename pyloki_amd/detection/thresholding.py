"""Monte-Carlo design and evaluation of EP threshold schemes on the GPU.

Mirrors ``pyloki.detection.thresholding`` (``src/pyloki/detection/thresholding.py``):
``simulate_folds``, ``prune_folds``, ``prune_folds_survival``, ``fold_max_scores``,
``Folds``, ``get_next_state``, ``gen_next_using_thresh``, ``gen_next_using_surv_prob``,
``transition_state``, ``DynamicThresholdScheme`` (legacy and improved modes),
``determine_scheme`` and ``evaluate_scheme``, with the same arguments, state records and
results.

Every simulated fold set is one task of the HIP kernel ``pyl_thresh_simulate``: it draws
the noise, forms the folds, normalises and scores them with the FFA path's boxcar scorer,
and either counts survivors for a list of thresholds or writes the survivors out. Within
a scheme run the folds of every state stay in HBM: a state's folds occupy a slot of
``2 * ntrials`` rows (a bound: ``simulate_folds`` yields ``n * ceil(ntrials / n) <
2 * ntrials`` rows from ``n <= 2 * ntrials - 1``), and the host sees only counts.

Random numbers. The reference draws ``numpy.random.Generator.normal`` from an unseeded
generator, so its runs are not reproducible. Here each simulated fold set has a 64-bit
*stream* and the run a 64-bit *seed* (``seed=`` keyword; ``None`` draws one from fresh OS
entropy); deviate (row i, bin j) of a set is a fixed function of (seed, stream, i, j)
(``pyl_rng_normal``). A fold set can therefore be simulated twice with the same result:
the DP counts survivors of every candidate transition first and materialises only the
winners' survivors afterwards. Streams of a run:

- ``evaluate_scheme`` / ``determine_scheme``: 0 (H0) and 1 (H1) for the initial folds,
  ``2 + 2 s + h`` for stage ``s``, hypothesis ``h``;
- ``DynamicThresholdScheme`` "improved": 0 / 1 initial, 2 / 3 the stage-0 simulation,
  ``4 + 2 (((s - 1) nthr + j) nprobs + k) + h`` for the simulation of previous state
  ``(j, k)`` at stage ``s >= 1``;
- "legacy": 0 / 1 initial, ``2 + 2 i + h`` for threshold ``i`` at stage 0,
  ``2 + 2 nthr + 2 ((((s - 1) nthr + i) nthr + j) nprobs + k) + h`` for the transition
  ``(j, k) -> i`` at stage ``s >= 1``.

The reference's functions are numba-compiled with the default (Python) error model:
``get_next_state`` raises ``ZeroDivisionError`` when the cumulative H1 success is zero,
and so do the runs that reach such a state; this module does the same.
"""

from __future__ import annotations

import ctypes
import math
from pathlib import Path
from typing import Literal

import numpy as np
import torch

from pyloki_amd import _lib
from pyloki_amd.detection.schemes import StateInfo, StatesInfo, bound_scheme, trials_scheme
from pyloki_amd.detection.scoring import generate_box_width_trials
from pyloki_amd.device import ptr, require_gpu, stream_ptr
from pyloki_amd.simulation.pulse import generate_folded_profile

TASK_DTYPE = np.dtype([
    ("src_row", np.int64), ("dst_row", np.int64), ("score_off", np.int64),
    ("count_off", np.int64), ("stream", np.uint64), ("var_norm", np.float64),
    ("thr", np.float64), ("n_prev", np.int32), ("n_sim", np.int32), ("hyp", np.int32),
    ("thr_first", np.int32), ("thr_count", np.int32), ("flags", np.int32)])

state_dtype = np.dtype([
    ("success_h0", np.float32), ("success_h1", np.float32), ("complexity", np.float32),
    ("complexity_cumul", np.float32), ("success_h1_cumul", np.float32),
    ("nbranches", np.float32), ("threshold", np.float32), ("cost", np.float32),
    ("threshold_prev", np.float32), ("success_h1_cumul_prev", np.float32),
    ("is_empty", np.bool_)])

_MAX_THR = 256


def neighbouring_indices(beam_indices: np.ndarray, target_idx: int, num: int) -> np.ndarray:
    """Up to ``num`` consecutive entries of the sorted ``beam_indices`` around
    ``target_idx`` (``thresholding.py:26-42``)."""
    beam_indices = np.asarray(beam_indices, dtype=np.int64)
    pos = int(np.searchsorted(beam_indices, target_idx))
    left = max(0, int(pos - num // 2))
    right = min(len(beam_indices), left + num)
    left = max(0, right - num)
    return beam_indices[left:right]


def _seed_stream(rng) -> tuple[int, int]:
    """(seed, stream) of one standalone simulation: a Generator draws a seed, an int is the
    seed itself, a (seed, stream) pair is used as given."""
    if isinstance(rng, np.random.Generator):
        return int(rng.integers(0, 2**63)), 0
    if isinstance(rng, tuple):
        return int(rng[0]), int(rng[1])
    return int(rng), 0


def _fresh_seed() -> int:
    return int(np.random.default_rng().integers(0, 2**63))


class _Sim:
    """Launches of ``pyl_thresh_simulate`` for one nbins / profile / width set."""

    def __init__(self, nbins: int, profile: np.ndarray, bias_snr: float, ducy_max: float,
                 wtsp: float, seed: int) -> None:
        self.dev = require_gpu()
        self.lib = _lib.load()
        self.nbins = int(nbins)
        self.widths = generate_box_width_trials(self.nbins, ducy_max=ducy_max, wtsp=wtsp)
        prof = np.zeros((2, self.nbins), dtype=np.float64)
        prof[1] = np.asarray(profile, dtype=np.float64) * float(bias_snr)
        self.prof = torch.from_numpy(prof).to(self.dev)
        self.seed = int(seed) & (2**64 - 1)

    def run(self, src: torch.Tensor | None, tasks: np.ndarray, mode: int, *,
            thr: torch.Tensor | None = None, dst: torch.Tensor | None = None,
            scores: torch.Tensor | None = None, counts: torch.Tensor,
            var_add: float = 1.0) -> None:
        if len(tasks) == 0:
            return
        t = torch.from_numpy(np.ascontiguousarray(tasks).view(np.uint8)).to(self.dev)
        w = self.widths
        _lib.check(self.lib.pyl_thresh_simulate(
            ptr(src), ptr(t), len(tasks), self.nbins, ptr(self.prof), float(var_add),
            ctypes.c_uint64(self.seed), w.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            len(w), ptr(thr), mode, ptr(dst), ptr(scores), ptr(counts), stream_ptr()))


def _n_sim(n_prev: int, ntrials_min: int) -> int:
    return int(n_prev) * int(math.ceil(ntrials_min / n_prev))


def _tasks(n: int) -> np.ndarray:
    t = np.zeros(n, dtype=TASK_DTYPE)
    t["src_row"] = -1
    return t


def _to_dev(arr: np.ndarray, dev) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float32)).to(dev)


# ---------------------------------------------------------------------- fold-level API
def simulate_folds(folds: np.ndarray, var_cur: float, profile: np.ndarray, rng,
                   bias_snr: float = 0, var_add: float = 1, ntrials_min: int = 1024
                   ) -> tuple[np.ndarray, float]:
    """``thresholding.py:45-95``: each of the ``n * ceil(ntrials_min / n)`` output rows is
    input row ``i % n`` plus noise of variance ``var_add`` plus ``profile * bias_snr``.
    ``rng``: a ``numpy.random.Generator`` (draws the seed), an int seed or a ``(seed,
    stream)`` pair."""
    folds = np.ascontiguousarray(folds, dtype=np.float32)
    n_prev, nbins = folds.shape
    if n_prev == 0:
        msg = "No trials in the input folds"
        raise ValueError(msg)
    seed, stream = _seed_stream(rng)
    sim = _Sim(nbins, profile, bias_snr, 0.2, 1.0, seed)
    n_sim = _n_sim(n_prev, ntrials_min)
    t = _tasks(1)
    t[0] = (0, 0, 0, 0, stream, var_cur + var_add, -np.inf, n_prev, n_sim, 1, 0, 0, 0)
    out = torch.empty((n_sim, nbins), dtype=torch.float32, device=sim.dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=sim.dev)
    sim.run(_to_dev(folds, sim.dev), t, _lib.PYL_SIM_COMPACT, dst=out, counts=cnt,
            var_add=var_add)
    return out.cpu().numpy(), var_cur + var_add


def fold_max_scores(folds: np.ndarray, var_cur: float, ducy_max: float = 0.2,
                    wtsp: float = 1.0) -> np.ndarray:
    """``thresholding.py:483-499``: per-trial maximum boxcar S/N of ``folds / sqrt(var)``
    (normalised in FP64, rounded to FP32)."""
    folds = np.ascontiguousarray(folds, dtype=np.float32)
    n, nbins = folds.shape
    if n == 0:
        return np.zeros(0, dtype=np.float32)
    sim = _Sim(nbins, np.zeros(nbins), 0.0, ducy_max, wtsp, 0)
    t = _tasks(1)
    t[0] = (0, 0, 0, 0, 0, var_cur, 0.0, n, n, 0, 0, 0, _lib.PYL_SIM_NO_NOISE)
    scores = torch.empty(n, dtype=torch.float32, device=sim.dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=sim.dev)
    thr = torch.zeros(1, dtype=torch.float64, device=sim.dev)
    sim.run(_to_dev(folds, sim.dev), t, _lib.PYL_SIM_COUNT, thr=thr, scores=scores, counts=cnt)
    return scores.cpu().numpy()


def prune_folds(folds: np.ndarray, var_cur: float, snr_thresh: float, ducy_max: float = 0.2,
                wtsp: float = 1.0) -> np.ndarray:
    """``thresholding.py:98-134``: the rows whose maximum S/N is above ``snr_thresh``, in
    order."""
    folds = np.ascontiguousarray(folds, dtype=np.float32)
    n, nbins = folds.shape
    if n == 0:
        return folds.copy()
    sim = _Sim(nbins, np.zeros(nbins), 0.0, ducy_max, wtsp, 0)
    t = _tasks(1)
    t[0] = (0, 0, 0, 0, 0, var_cur, snr_thresh, n, n, 0, 0, 0, _lib.PYL_SIM_NO_NOISE)
    out = torch.empty((n, nbins), dtype=torch.float32, device=sim.dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=sim.dev)
    sim.run(_to_dev(folds, sim.dev), t, _lib.PYL_SIM_COMPACT, dst=out, counts=cnt)
    return out[: int(cnt.item())].cpu().numpy()


def prune_folds_survival(folds: np.ndarray, var_cur: float, survive_prob: float,
                         ducy_max: float = 0.2, wtsp: float = 1.0
                         ) -> tuple[float, np.ndarray]:
    """``thresholding.py:137-157``: the best ``int(survive_prob * n)`` rows by score, best
    first, and the score of the last one kept (the reference's unstable argsort, reversed;
    equal scores are kept in reverse row order here)."""
    folds = np.ascontiguousarray(folds, dtype=np.float32)
    scores = fold_max_scores(folds, var_cur, ducy_max, wtsp)
    n_surv = int(survive_prob * len(scores))
    order = np.argsort(scores, kind="stable")[::-1][:n_surv]
    threshold = float(scores[order[-1]])
    return threshold, np.ascontiguousarray(folds[order])


class Folds:
    """Surviving H0 / H1 folds of a state (``thresholding.py:160-243``): ``Folds(h0, h1,
    variance)`` or ``Folds(h0, h1, scores_h0, scores_h1, variance)``."""

    def __init__(self, folds_h0: np.ndarray, folds_h1: np.ndarray, third, fourth=None,
                 fifth=None) -> None:
        if fourth is None:
            if not isinstance(third, (float, np.floating)):
                msg = "Legacy Folds expects variance as the third argument"
                raise TypeError(msg)
            scores_h0 = np.empty(0, dtype=np.float32)
            scores_h1 = np.empty(0, dtype=np.float32)
            variance = float(third)
        else:
            if fifth is None:
                msg = "Scored Folds requires variance as the fifth argument"
                raise TypeError(msg)
            scores_h0, scores_h1, variance = third, fourth, float(fifth)
        self.folds_h0 = np.ascontiguousarray(folds_h0, dtype=np.float32)
        self.folds_h1 = np.ascontiguousarray(folds_h1, dtype=np.float32)
        self.scores_h0 = np.ascontiguousarray(scores_h0, dtype=np.float32)
        self.scores_h1 = np.ascontiguousarray(scores_h1, dtype=np.float32)
        self.variance = variance
        self.is_empty = 0 if (len(self.folds_h0) > 0 and len(self.folds_h1) > 0) else 1


def create_empty_folds() -> Folds:
    e = np.empty((0, 1), dtype=np.float32)
    z = np.empty(0, dtype=np.float32)
    return Folds(e, e, z, z, 0.0)


def _next_states(cur: np.ndarray, threshold, success_h0, success_h1, nbranches) -> np.ndarray:
    """Vectorised ``get_next_state`` (``thresholding.py:350-376``): FP64 locals from the
    FP32 fields of ``cur``, stored as FP32."""
    cur = np.asarray(cur)
    s0 = np.asarray(success_h0, dtype=np.float64)
    s1 = np.asarray(success_h1, dtype=np.float64)
    # complexity * nbranches and complexity_cumul + that are FP32 operations when the
    # branching pattern is FP32 (numba types them so), FP64 otherwise.
    nbr = np.asarray(nbranches)
    ft = np.float32 if nbr.dtype == np.float32 else np.float64
    nbr = nbr.astype(ft)
    nleaves_next = cur["complexity"].astype(ft) * nbr
    nleaves_surv = nleaves_next.astype(np.float64) * s0
    cc = cur["complexity_cumul"].astype(ft) + nleaves_next
    s1c = cur["success_h1_cumul"].astype(np.float64) * s1
    if np.any(s1c == 0.0):
        msg = "division by zero"
        raise ZeroDivisionError(msg)
    shape = np.broadcast(nleaves_surv, np.asarray(threshold)).shape
    out = np.zeros(shape, dtype=state_dtype)
    out["success_h0"] = s0
    out["success_h1"] = s1
    out["complexity"] = nleaves_surv
    out["complexity_cumul"] = cc
    out["success_h1_cumul"] = s1c
    out["nbranches"] = nbr
    out["threshold"] = threshold
    out["cost"] = cc.astype(np.float64) / s1c
    out["threshold_prev"] = cur["threshold"]
    out["success_h1_cumul_prev"] = cur["success_h1_cumul"]
    return out


def get_next_state(state_cur: np.ndarray, threshold: float, success_h0: float,
                   success_h1: float, nbranches: float) -> np.ndarray:
    """``thresholding.py:350-376``: the state after one stage (a one-record array)."""
    return _next_states(np.asarray(state_cur)[()], threshold, success_h0, success_h1,
                        nbranches).reshape(1)


def _initial_state() -> np.ndarray:
    s = np.ones(1, dtype=state_dtype)[0]
    s["threshold"] = -1
    s["threshold_prev"] = -1
    return s


def gen_next_using_thresh(state_cur, folds_cur: Folds, threshold: float, nbranches: float,
                          bias_snr: float, profile: np.ndarray, rng, ntrials: int = 1024,
                          ducy_max: float = 0.2, wtsp: float = 1.0
                          ) -> tuple[np.ndarray, Folds]:
    """``thresholding.py:379-428``: simulate H0 and H1 one stage on, keep ``score >
    threshold``."""
    folds_h0, variance = simulate_folds(folds_cur.folds_h0, folds_cur.variance, profile, rng,
                                        bias_snr=0, var_add=1, ntrials_min=ntrials)
    h0 = prune_folds(folds_h0, variance, threshold, ducy_max=ducy_max, wtsp=wtsp)
    folds_h1, variance = simulate_folds(folds_cur.folds_h1, folds_cur.variance, profile, rng,
                                        bias_snr=bias_snr, var_add=1, ntrials_min=ntrials)
    h1 = prune_folds(folds_h1, variance, threshold, ducy_max=ducy_max, wtsp=wtsp)
    state = get_next_state(state_cur, threshold, len(h0) / len(folds_h0),
                           len(h1) / len(folds_h1), nbranches)
    return state, Folds(h0, h1, variance)


def gen_next_using_surv_prob(state_cur, folds_cur: Folds, surv_prob_h0: float,
                             nbranches: float, bias_snr: float, profile: np.ndarray, rng,
                             ntrials: int = 1024, ducy_max: float = 0.2, wtsp: float = 1.0
                             ) -> tuple[np.ndarray, Folds]:
    """``thresholding.py:431-480``: the H0 threshold that keeps ``surv_prob_h0`` of the
    H0 trials, applied to H1."""
    folds_h0, variance = simulate_folds(folds_cur.folds_h0, folds_cur.variance, profile, rng,
                                        bias_snr=0, var_add=1, ntrials_min=ntrials)
    thr, h0 = prune_folds_survival(folds_h0, variance, surv_prob_h0, ducy_max=ducy_max,
                                   wtsp=wtsp)
    folds_h1, variance = simulate_folds(folds_cur.folds_h1, folds_cur.variance, profile, rng,
                                        bias_snr=bias_snr, var_add=1, ntrials_min=ntrials)
    h1 = prune_folds(folds_h1, variance, thr, ducy_max=ducy_max, wtsp=wtsp)
    state = get_next_state(state_cur, thr, len(h0) / len(folds_h0), len(h1) / len(folds_h1),
                           nbranches)
    return state, Folds(h0, h1, variance)


def transition_state(state_cur, folds_cur: Folds, threshold: float, nbranches: float
                     ) -> tuple[np.ndarray, Folds]:
    """``thresholding.py:502-525``: keep the scored folds above ``threshold``."""
    k0 = np.nonzero(folds_cur.scores_h0 > threshold)[0]
    k1 = np.nonzero(folds_cur.scores_h1 > threshold)[0]
    state = get_next_state(state_cur, threshold, len(k0) / len(folds_cur.folds_h0),
                           len(k1) / len(folds_cur.folds_h1), nbranches)
    return state, Folds(folds_cur.folds_h0[k0], folds_cur.folds_h1[k1],
                        folds_cur.scores_h0[k0], folds_cur.scores_h1[k1], folds_cur.variance)


# ------------------------------------------------------------------ single-path schemes
class _Path:
    """The H0 / H1 folds of one scheme path, resident in HBM between stages."""

    def __init__(self, sim: _Sim, ntrials: int) -> None:
        self.sim_ = sim
        self.ntrials = int(ntrials)
        self.cap = 2 * self.ntrials
        dev = sim.dev
        self.buf = [torch.empty((2, self.cap, sim.nbins), dtype=torch.float32, device=dev)
                    for _ in range(2)]
        self.scores = torch.empty((2, self.cap), dtype=torch.float32, device=dev)
        self.counts = torch.zeros(2, dtype=torch.int32, device=dev)
        self.cur = 0
        self.n = [self.ntrials, self.ntrials]

    def initial(self, var_init: float) -> None:
        """``simulate_folds`` of ``ntrials`` zero rows, streams 0 (H0) and 1 (H1)."""
        t = _tasks(2)
        for h in range(2):
            t[h] = (-1, h * self.cap, 0, h, h, var_init, -np.inf, self.ntrials, self.ntrials, h,
                    0, 0, 0)
        self.sim_.run(None, t, _lib.PYL_SIM_COMPACT, dst=self.buf[self.cur],
                      counts=self.counts, var_add=var_init)

    def sim(self, stage: int, var_cur: float, thr: dict[int, float]) -> tuple[list, list]:
        """Simulate the hypotheses in ``thr`` one stage on (streams ``2 + 2 stage + h``)
        into the other buffer, keeping ``score > thr[h]`` (scores of the kept rows too);
        returns (simulated rows, kept rows) per hypothesis (0 for one not simulated)."""
        hs = sorted(thr)
        t = _tasks(len(hs))
        n_sim = [0, 0]
        for c, h in enumerate(hs):
            n_sim[h] = _n_sim(self.n[h], self.ntrials)
            t[c] = (h * self.cap, h * self.cap, h * self.cap, h, 2 + 2 * stage + h,
                    var_cur + 1.0, thr[h], self.n[h], n_sim[h], h, 0, 0, 0)
        self.sim_.run(self.buf[self.cur].view(-1, self.sim_.nbins), t, _lib.PYL_SIM_COMPACT,
                      dst=self.buf[1 - self.cur], scores=self.scores, counts=self.counts)
        got = self.counts.cpu().numpy().astype(int).tolist()
        return n_sim, [got[h] if h in thr else 0 for h in range(2)]

    def swap(self, n: list[int]) -> None:
        self.cur = 1 - self.cur
        self.n = list(n)


def evaluate_scheme(thresholds: np.ndarray, branching_pattern: np.ndarray, ref_ducy: float,
                    nbins: int = 64, ntrials: int = 2048, snr_final: float = 8,
                    ducy_max: float = 0.2, wtsp: float = 1.0, *, seed: int | None = None
                    ) -> StatesInfo:
    """``thresholding.py:1062-1110``: the states of a given threshold scheme (H0 false
    alarms and H1 survival, stage by stage) from Monte-Carlo folds."""
    var_init = 1.0
    branching_pattern = np.asarray(branching_pattern)
    thresholds = np.asarray(thresholds)
    nstages = len(branching_pattern)
    profile = generate_folded_profile(nbins=nbins, ducy=ref_ducy)
    bias_snr = snr_final / np.sqrt(nstages + 1)
    if len(thresholds) != nstages:
        msg = "Number of thresholds must match the number of stages"
        raise ValueError(msg)
    sim = _Sim(nbins, profile, bias_snr, ducy_max, wtsp, _fresh_seed() if seed is None else seed)
    path = _Path(sim, ntrials)
    path.initial(var_init)
    states: list[np.ndarray] = []
    prev = _initial_state()
    var_cur = var_init
    for istage in range(nstages):
        if istage > 0 and min(path.n) == 0:
            break  # "Path not viable, No trials survived, stopping"
        thr = float(thresholds[istage])
        n_sim, kept = path.sim(istage, var_cur, {0: thr, 1: thr})
        path.swap(kept)
        var_cur += 1.0
        cur = get_next_state(prev, thresholds[istage], kept[0] / n_sim[0], kept[1] / n_sim[1],
                             branching_pattern[istage])[0]
        states.append(cur)
        prev = cur
    return StatesInfo([StateInfo.from_record(s) for s in states])


def determine_scheme(survive_probs: np.ndarray, branching_pattern: np.ndarray,
                     ref_ducy: float, nbins: int = 64, ntrials: int = 2048,
                     snr_final: float = 8, ducy_max: float = 0.2, wtsp: float = 1.0, *,
                     seed: int | None = None) -> StatesInfo:
    """``thresholding.py:1010-1059``: the scheme whose stage thresholds keep the given H0
    survival fractions."""
    survive_probs = np.asarray(survive_probs, dtype=np.float64)
    branching_pattern = np.asarray(branching_pattern)
    if len(survive_probs) != len(branching_pattern):
        msg = "Number of survive_probs must match the number of stages"
        raise ValueError(msg)
    var_init = 1.0
    nstages = len(branching_pattern)
    profile = generate_folded_profile(nbins=nbins, ducy=ref_ducy)
    bias_snr = snr_final / np.sqrt(nstages + 1)
    sim = _Sim(nbins, profile, bias_snr, ducy_max, wtsp, _fresh_seed() if seed is None else seed)
    path = _Path(sim, ntrials)
    path.initial(var_init)
    lib = _lib.load()
    states: list[np.ndarray] = []
    prev = _initial_state()
    var_cur = var_init
    for istage in range(nstages):
        if istage > 0 and min(path.n) == 0:
            break
        # H0: simulate and score every row, keep the best int(p * n) (best first); H1:
        # the same simulation again, keeping the rows above the H0 threshold.
        n_sim, _ = path.sim(istage, var_cur, {0: -np.inf})
        sc0 = path.scores[0, : n_sim[0]].cpu().numpy()
        n_surv = int(survive_probs[istage] * n_sim[0])
        order = np.argsort(sc0, kind="stable")[::-1][:n_surv]
        thr = float(sc0[order[-1]])
        out = path.buf[1 - path.cur]
        tmp = torch.empty_like(out[0])
        idx = torch.from_numpy(np.ascontiguousarray(order, dtype=np.int64)).to(sim.dev)
        _lib.check(lib.pyl_rows_gather(ptr(out[0]), ptr(idx), len(order), sim.nbins, ptr(tmp),
                                       stream_ptr()))
        out[0].copy_(tmp)
        n_sim1, kept = path.sim(istage, var_cur, {1: thr})
        n_sim[1] = n_sim1[1]
        path.swap([len(order), kept[1]])
        var_cur += 1.0
        cur = get_next_state(prev, thr, path.n[0] / n_sim[0], path.n[1] / n_sim[1],
                             branching_pattern[istage])[0]
        states.append(cur)
        prev = cur
    return StatesInfo([StateInfo.from_record(s) for s in states])


# ------------------------------------------------------------------ dynamic programming
class DynamicThresholdScheme:
    """Beam-searched dynamic program over (stage, threshold, H1-success bin)
    (``thresholding.py:702-1007``): per stage, every state on the beam is simulated one
    stage on and moved to every nearby threshold of the next beam; of the transitions that
    land in one (threshold, success bin) cell the one with the least cumulative H0
    complexity is kept (the first such in the reference's loop order on ties).

    ``mode="improved"`` simulates each previous state once per stage and applies every
    threshold to that simulation; ``"legacy"`` simulates every transition separately. The
    folds of all states stay on the GPU; ``run`` moves only survivor counts to the host.
    ``seed`` fixes the run's random numbers (``None``: fresh entropy, as the reference).
    """

    def __init__(self, branching_pattern: np.ndarray, ref_ducy: float, nbins: int = 64,
                 ntrials: int = 1024, nprobs: int = 10, prob_min: float = 0.05,
                 snr_final: float = 8, nthresholds: int = 100, ducy_max: float = 0.2,
                 wtsp: float = 1.0, beam_width: float = 0.7,
                 mode: Literal["legacy", "improved"] = "legacy", *,
                 seed: int | None = None) -> None:
        if mode not in ("legacy", "improved"):
            msg = f"mode must be 'legacy' or 'improved', got {mode!r}"
            raise ValueError(msg)
        self.mode = mode
        self.seed = _fresh_seed() if seed is None else int(seed)
        self.branching_pattern = np.asarray(branching_pattern)
        self.ref_ducy = ref_ducy
        self.profile = generate_folded_profile(nbins=nbins, ducy=ref_ducy)
        self.ntrials = int(ntrials)
        self.snr_final = snr_final
        self.ducy_max = ducy_max
        self.wtsp = wtsp
        self.beam_width = beam_width
        self.thresholds = np.linspace(0.1, self.snr_final, nthresholds)
        if self.nthresholds > _MAX_THR:
            msg = f"at most {_MAX_THR} thresholds are supported, got {self.nthresholds}"
            raise ValueError(msg)
        self.prob_min = prob_min
        self.probs = np.logspace(np.log10(prob_min), 0, nprobs)
        state = np.ones(1, dtype=state_dtype)
        state["threshold"] = -1.0
        state["threshold_prev"] = -1.0
        self.states = np.full((self.nstages, self.nthresholds, self.nprobs), state)
        self.guess_path = np.minimum(bound_scheme(self.nstages, self.snr_final),
                                     trials_scheme(self.branching_pattern, trials_start=1))
        self._sim = _Sim(self.nbins, self.profile, self.bias_snr, ducy_max, wtsp, self.seed)
        self._cap = 2 * self.ntrials
        nslots = self.nthresholds * self.nprobs * 2
        dev = self._sim.dev
        self._bank = [torch.empty((nslots * self._cap, self.nbins), dtype=torch.float32,
                                  device=dev) for _ in range(2)]
        self._in = 0
        # Rows of every state's H0 / H1 folds and their variance (the input bank).
        self._nrows = np.zeros((self.nthresholds, self.nprobs, 2), dtype=np.int64)
        self._var = np.zeros((self.nthresholds, self.nprobs), dtype=np.float64)
        self._thr_dev = torch.from_numpy(self.thresholds.astype(np.float64)).to(dev)
        if self.mode == "improved":
            self._init_improved()
        else:
            self._init_legacy()

    @property
    def nstages(self) -> int:
        return len(self.branching_pattern)

    @property
    def bias_snr(self) -> float:
        return self.snr_final / np.sqrt(self.nstages + 1)

    @property
    def nthresholds(self) -> int:
        return len(self.thresholds)

    @property
    def nprobs(self) -> int:
        return len(self.probs)

    @property
    def nbins(self) -> int:
        return len(self.profile)

    def get_current_thresholds_idx(self, istage: int) -> np.ndarray:
        guess = self.guess_path[istage]
        lower = max(0, guess - self.beam_width)
        upper = min(self.snr_final, guess + self.beam_width)
        return np.where((self.thresholds >= lower) & (self.thresholds <= upper))[0]

    # -- device helpers
    def _slot(self, i, k, h):
        return ((np.asarray(i) * self.nprobs + np.asarray(k)) * 2 + np.asarray(h)) * self._cap

    def _initial_folds(self, var_init: float) -> torch.Tensor:
        """The H0 / H1 folds simulated from ``ntrials`` zero rows (streams 0 and 1)."""
        buf = torch.empty((2 * self._cap, self.nbins), dtype=torch.float32, device=self._sim.dev)
        t = _tasks(2)
        for h in range(2):
            t[h] = (-1, h * self._cap, 0, h, h, var_init, -np.inf, self.ntrials, self.ntrials, h,
                    0, 0, 0)
        cnt = torch.zeros(2, dtype=torch.int32, device=self._sim.dev)
        self._sim.run(None, t, _lib.PYL_SIM_COMPACT, dst=buf, counts=cnt, var_add=var_init)
        return buf

    def _count(self, src: torch.Tensor, tasks: np.ndarray) -> np.ndarray:
        counts = torch.zeros(int(tasks["count_off"].max()) + int(tasks["thr_count"].max()),
                             dtype=torch.int32, device=self._sim.dev)
        self._sim.run(src, tasks, _lib.PYL_SIM_COUNT, thr=self._thr_dev, counts=counts)
        return counts.cpu().numpy()

    def _materialize(self, src: torch.Tensor, tasks: np.ndarray, dst: torch.Tensor) -> np.ndarray:
        counts = torch.zeros(len(tasks), dtype=torch.int32, device=self._sim.dev)
        self._sim.run(src, tasks, _lib.PYL_SIM_COMPACT, dst=dst, counts=counts)
        return counts.cpu().numpy()

    def _select(self, stage: int, cand_i: np.ndarray, prev: np.ndarray, s0: np.ndarray,
                s1: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """New states of candidate transitions (in the reference's loop order) and, for
        every (threshold, success bin) cell, the index of the transition kept there."""
        st = _next_states(prev, self.thresholds[cand_i], s0, s1, self.branching_pattern[stage])
        iprob = np.digitize(st["success_h1_cumul"], self.probs) - 1
        ok = (iprob >= 0) & (iprob < self.nprobs)
        idx = np.nonzero(ok)[0]
        order = np.lexsort((idx, st["complexity_cumul"][idx], iprob[idx], cand_i[idx]))
        idx = idx[order]
        cell = cand_i[idx] * self.nprobs + iprob[idx]
        first = np.ones(len(idx), dtype=bool)
        first[1:] = cell[1:] != cell[:-1]
        win = idx[first]
        return st, np.stack([win, iprob[win]], axis=1) if len(win) else np.zeros((0, 2), int)

    def _store(self, stage: int, st: np.ndarray, cand_i: np.ndarray, winners: np.ndarray,
               src: torch.Tensor, sim_tasks: np.ndarray, var_new: np.ndarray) -> None:
        """Write the winners' states and materialise their survivors into the output bank.
        ``sim_tasks[w, h]`` is the simulation task of winner candidate ``w``."""
        out_bank = self._bank[1 - self._in]
        nrows = np.zeros_like(self._nrows)
        var = np.zeros_like(self._var)
        if len(winners):
            w, ip = winners[:, 0], winners[:, 1]
            i = cand_i[w]
            self.states[stage, i, ip] = st[w]
            t = np.concatenate([sim_tasks[w, 0], sim_tasks[w, 1]]).copy()
            h = np.repeat([0, 1], len(w))
            t["dst_row"] = self._slot(np.tile(i, 2), np.tile(ip, 2), h)
            t["thr"] = np.tile(self.thresholds[i], 2)
            t["count_off"] = np.arange(len(t))
            t["thr_count"] = 0
            kept = self._materialize(src, t, out_bank)
            nrows[i, ip, 0] = kept[: len(w)]
            nrows[i, ip, 1] = kept[len(w):]
            var[i, ip] = var_new[w]
        self._nrows = nrows
        self._var = var
        self._in = 1 - self._in

    # -- stage 0
    def _init_improved(self) -> None:
        var_init = 1.0
        buf = self._initial_folds(var_init)
        t = _tasks(2)
        for h in range(2):
            t[h] = (h * self._cap, 0, 0, h * self.nthresholds, 2 + h, var_init + 1.0, 0.0,
                    self.ntrials, self.ntrials, h, 0, self.nthresholds, 0)
        cnt = self._count(buf, t).reshape(2, self.nthresholds)
        beam = self.get_current_thresholds_idx(0)
        prev = np.full(len(beam), _initial_state())
        s0 = cnt[0, beam] / self.ntrials
        s1 = cnt[1, beam] / self.ntrials
        st, win = self._select(0, beam, prev, s0, s1)
        sim_tasks = np.stack([np.repeat(t[0:1], len(beam)), np.repeat(t[1:2], len(beam))], axis=1)
        self._in = 1  # the winners go to bank 0
        self._store(0, st, beam, win, buf, sim_tasks, np.full(len(beam), var_init + 1.0))

    def _init_legacy(self) -> None:
        var_init = 1.0
        buf = self._initial_folds(var_init)
        beam = self.get_current_thresholds_idx(0)
        nb = len(beam)
        t = _tasks(2 * nb)
        for c, i in enumerate(beam.tolist()):
            for h in range(2):
                t[2 * c + h] = (h * self._cap, 0, 0, 2 * c + h, 2 + 2 * i + h, var_init + 1.0, 0.0,
                                self.ntrials, self.ntrials, h, i, 1, 0)
        cnt = self._count(buf, t).reshape(nb, 2)
        prev = np.full(nb, _initial_state())
        st, win = self._select(0, beam, prev, cnt[:, 0] / self.ntrials, cnt[:, 1] / self.ntrials)
        self._in = 1
        self._store(0, st, beam, win, buf, t.reshape(nb, 2), np.full(nb, var_init + 1.0))

    # -- stages 1 ..
    def _candidates(self, istage: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(i, j, k) of every transition the reference evaluates at ``istage``, in its loop
        order, restricted to previous states that exist and have folds."""
        beam_cur = self.get_current_thresholds_idx(istage)
        beam_prev = self.get_current_thresholds_idx(istage - 1)
        prev_ok = ~self.states[istage - 1]["is_empty"] & np.all(self._nrows > 0, axis=2)
        ci, cj, ck = [], [], []
        for i in beam_cur.tolist():
            for j in neighbouring_indices(beam_prev, i, self._thres_neigh).tolist():
                ks = np.nonzero(prev_ok[j])[0]
                ci.extend([i] * len(ks))
                cj.extend([j] * len(ks))
                ck.extend(ks.tolist())
        return (np.asarray(ci, dtype=np.int64), np.asarray(cj, dtype=np.int64),
                np.asarray(ck, dtype=np.int64))

    def _stage(self, istage: int) -> None:
        ci, cj, ck = self._candidates(istage)
        src = self._bank[self._in]
        n = self._nrows
        prev = self.states[istage - 1][cj, ck]
        if self.mode == "improved":
            pairs = np.unique(cj * self.nprobs + ck)
            pj, pk = pairs // self.nprobs, pairs % self.nprobs
            t = _tasks(2 * len(pairs))
            for h in range(2):
                sl = slice(h, None, 2)
                n_prev = n[pj, pk, h]
                t["src_row"][sl] = self._slot(pj, pk, h)
                t["count_off"][sl] = (np.arange(len(pairs)) * 2 + h) * self.nthresholds
                t["stream"][sl] = (4 + 2 * (((istage - 1) * self.nthresholds + pj) * self.nprobs
                                             + pk) + h).astype(np.uint64)
                t["var_norm"][sl] = self._var[pj, pk] + 1.0
                t["n_prev"][sl] = n_prev
                t["n_sim"][sl] = n_prev * np.ceil(self.ntrials / n_prev).astype(np.int64)
                t["hyp"][sl] = h
                t["thr_count"][sl] = self.nthresholds
            cnt = self._count(src, t).reshape(len(pairs), 2, self.nthresholds) if len(t) else None
            pos = np.searchsorted(pairs, cj * self.nprobs + ck)
            c0 = cnt[pos, 0, ci] if len(ci) else np.zeros(0)
            c1 = cnt[pos, 1, ci] if len(ci) else np.zeros(0)
            s0 = c0 / t["n_sim"][2 * pos]
            s1 = c1 / t["n_sim"][2 * pos + 1]
            sim_tasks = t.reshape(len(pairs), 2)[pos]
        else:
            nc = len(ci)
            t = _tasks(2 * nc)
            for h in range(2):
                sl = slice(h, None, 2)
                n_prev = n[cj, ck, h]
                t["src_row"][sl] = self._slot(cj, ck, h)
                t["count_off"][sl] = np.arange(nc) * 2 + h
                t["stream"][sl] = (2 + 2 * self.nthresholds + 2 * ((((istage - 1) * self.nthresholds
                                                                    + ci) * self.nthresholds + cj)
                                                                   * self.nprobs + ck) + h
                                   ).astype(np.uint64)
                t["var_norm"][sl] = self._var[cj, ck] + 1.0
                t["n_prev"][sl] = n_prev
                t["n_sim"][sl] = n_prev * np.ceil(self.ntrials / n_prev).astype(np.int64)
                t["hyp"][sl] = h
                t["thr_first"][sl] = ci
                t["thr_count"][sl] = 1
            cnt = self._count(src, t).reshape(nc, 2) if nc else np.zeros((0, 2))
            s0 = cnt[:, 0] / t["n_sim"][0::2]
            s1 = cnt[:, 1] / t["n_sim"][1::2]
            sim_tasks = t.reshape(nc, 2)
        st, win = self._select(istage, ci, prev, s0, s1)
        self._store(istage, st, ci, win, src, sim_tasks, self._var[cj, ck] + 1.0)

    def run(self, thres_neigh: int = 11) -> None:
        """Stages 1 .. nstages-1 (``thresholding.py:904-962``)."""
        self._thres_neigh = int(thres_neigh)
        for istage in range(1, self.nstages):
            self._stage(istage)

    def save(self, outdir: str = ".") -> str:
        """The scheme (attributes, arrays and the state table) as an ``.npz`` archive
        (``thresholding.py:965-1007`` writes the same members to HDF5)."""
        filebase = (
            f"dynscheme_nstages_{self.nstages:03d}_"
            f"nthresh_{self.nthresholds:03d}_nprobs_{self.nprobs:03d}_"
            f"ntrials_{self.ntrials:04d}_snr_{self.snr_final:04.1f}_"
            f"ducy_{self.ref_ducy:04.2f}_beam_{self.beam_width:03.1f}"
        )
        filename = Path(outdir) / f"{filebase}.npz"
        members = {f"attrs/{a}": np.array(getattr(self, a))
                   for a in ("ntrials", "snr_final", "ref_ducy", "ducy_max", "wtsp",
                             "beam_width", "mode")}
        for a in ("branching_pattern", "profile", "thresholds", "probs", "guess_path",
                  "states"):
            members[a] = np.asarray(getattr(self, a))
        np.savez_compressed(filename, **members)
        return filename.as_posix()
