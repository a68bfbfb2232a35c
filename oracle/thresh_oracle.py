"""CPU oracle of the Monte-Carlo threshold design -- TEST INFRASTRUCTURE ONLY.

Restates ``detection/thresholding.py`` (simulate_folds :45-95, prune_folds :98-134,
prune_folds_survival :137-157, get_next_state :350-376, transition_state :502-525,
fold_max_scores :483-499, the DP of DynamicThresholdScheme :702-962 in both modes,
determine_scheme :1010-1059, evaluate_scheme :1062-1110) with the reference's own loops,
on the C oracle's generator and scorer (``pyl_oracle.c``: ``orc_simulate_folds``,
``orc_boxcar_snr_2d``). The random streams follow the layout the product documents
(pyloki_amd/detection/thresholding.py), so a seeded GPU run can be compared bit for bit.

Pinning (tests/test_thresh_oracle.py): Philox4x32-10 against the Random123 known-answer
vectors; ln / sin / cos against libm within a few ulp; the normal deviates' moments and
a Kolmogorov-Smirnov test; simulate_folds against its closed form with zero noise;
get_next_state against exact rational arithmetic; the DP's cell selection against a
brute-force minimum; the scorer is the one pinned by the boxcar identities.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np

from oracle import oracle as _orc

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u64 = ctypes.c_uint64
_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_bound = False

STATE_DTYPE = np.dtype([
    ("success_h0", np.float32), ("success_h1", np.float32), ("complexity", np.float32),
    ("complexity_cumul", np.float32), ("success_h1_cumul", np.float32),
    ("nbranches", np.float32), ("threshold", np.float32), ("cost", np.float32),
    ("threshold_prev", np.float32), ("success_h1_cumul_prev", np.float32),
    ("is_empty", np.bool_)])


def _lib():
    global _bound
    L = _orc.lib()
    if not _bound:
        L.orc_philox4x32_10.restype = None
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_ln_det.restype = _dbl
        L.orc_ln_det.argtypes = [_dbl]
        L.orc_sincos_2pi.restype = None
        L.orc_sincos_2pi.argtypes = [_dbl, ctypes.POINTER(_dbl), ctypes.POINTER(_dbl)]
        L.orc_normal4.restype = None
        L.orc_normal4.argtypes = [_u64, _u64, ctypes.c_uint32, ctypes.c_uint32, _f64p]
        L.orc_rng_normal.restype = None
        L.orc_rng_normal.argtypes = [_u64, _u64, _i64, _i64, _dbl, _f32p]
        L.orc_simulate_folds.restype = None
        L.orc_simulate_folds.argtypes = [ctypes.c_void_p, _i64, _i64, _i64, _f64p, _dbl, _u64,
                                         _u64, _f32p]
        _bound = True
    return L


def philox(ctr, key) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint32)
    _lib().orc_philox4x32_10(np.asarray(ctr, dtype=np.uint32), np.asarray(key, dtype=np.uint32),
                             out)
    return out


def ln_det(x: float) -> float:
    return float(_lib().orc_ln_det(x))


def sincos_2pi(u: float) -> tuple[float, float]:
    s, c = _dbl(), _dbl()
    _lib().orc_sincos_2pi(u, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def rng_normal(seed: int, stream: int, n_rows: int, row_len: int, std: float = 1.0) -> np.ndarray:
    out = np.empty((n_rows, row_len), dtype=np.float32)
    _lib().orc_rng_normal(seed, stream, n_rows, row_len, std, out)
    return out


def simulate_folds(folds, var_cur, prof_scaled, seed, stream, var_add=1.0, ntrials_min=1024):
    """thresholding.py:45-95 on the shared generator; ``folds`` None = ntrials_min zero rows
    (the initial simulation of a scheme)."""
    nbins = len(prof_scaled)
    if folds is None:
        n_prev, src = ntrials_min, None
    else:
        folds = np.ascontiguousarray(folds, dtype=np.float32)
        n_prev = len(folds)
        if n_prev == 0:
            raise ValueError("No trials in the input folds")
        src = folds.ctypes.data_as(ctypes.c_void_p)
    n_sim = n_prev * int(math.ceil(ntrials_min / n_prev))
    out = np.empty((n_sim, nbins), dtype=np.float32)
    _lib().orc_simulate_folds(src, n_prev, n_sim, nbins,
                              np.ascontiguousarray(prof_scaled, dtype=np.float64), var_add,
                              seed, stream, out)
    return out, var_cur + var_add


def widths(nbins, ducy_max, wtsp):
    """scoring.py:10-39."""
    wmax = int(max(1, ducy_max * nbins))
    w = [1]
    while w[-1] <= wmax:
        nw = int(max(w[-1] + 1, int(wtsp * w[-1])))
        if nw > wmax:
            break
        w.append(nw)
    return np.asarray(w, dtype=np.int64)


def fold_max_scores(folds, var_cur, ducy_max=0.2, wtsp=1.0):
    """thresholding.py:483-499: max over widths of the boxcar S/N of f32(folds / sqrt(var))."""
    folds = np.ascontiguousarray(folds, dtype=np.float32)
    if len(folds) == 0:
        return np.zeros(0, dtype=np.float32)
    norm = (folds.astype(np.float64) / np.sqrt(np.float64(var_cur))).astype(np.float32)
    return _orc.boxcar_snr(norm, widths(folds.shape[1], ducy_max, wtsp), 1.0).max(axis=1)


def prune_folds(folds, var_cur, thr, ducy_max=0.2, wtsp=1.0):
    sc = fold_max_scores(folds, var_cur, ducy_max, wtsp)
    return np.ascontiguousarray(folds[np.nonzero(sc > thr)[0]])


def get_next_state(cur, threshold, s0, s1, nbranches):
    """thresholding.py:350-376, scalar, numba's type promotion."""
    nbr = np.float32(nbranches) if isinstance(nbranches, np.float32) else float(nbranches)
    if isinstance(nbr, np.float32):
        nleaves_next = np.float32(cur["complexity"]) * nbr
        cc = np.float32(cur["complexity_cumul"]) + nleaves_next
    else:
        nleaves_next = float(cur["complexity"]) * nbr
        cc = float(cur["complexity_cumul"]) + nleaves_next
    nleaves_surv = float(nleaves_next) * s0
    s1c = float(cur["success_h1_cumul"]) * s1
    if s1c == 0.0:
        raise ZeroDivisionError("division by zero")
    out = np.zeros(1, dtype=STATE_DTYPE)[0]
    out["success_h0"] = s0
    out["success_h1"] = s1
    out["complexity"] = nleaves_surv
    out["complexity_cumul"] = cc
    out["success_h1_cumul"] = s1c
    out["nbranches"] = nbranches
    out["threshold"] = threshold
    out["cost"] = float(cc) / s1c
    out["threshold_prev"] = cur["threshold"]
    out["success_h1_cumul_prev"] = cur["success_h1_cumul"]
    return out


def initial_state():
    s = np.ones(1, dtype=STATE_DTYPE)[0]
    s["threshold"] = -1
    s["threshold_prev"] = -1
    return s


def folded_profile(nbins, ducy, center=0.5):
    """simulation/pulse.py:512-523."""
    phase = np.linspace(0, 1, nbins, endpoint=False)
    sigma = ducy / (2 * np.sqrt(2 * np.log(10)))
    wrapped = np.mod(phase - center + 0.5, 1) - 0.5
    p = np.exp(-(wrapped**2) / (2 * sigma**2))
    p = p / p.max()
    return p / np.sqrt(np.sum(p**2))


def _profiles(nbins, ref_ducy, bias):
    prof = folded_profile(nbins, ref_ducy)
    return np.zeros(nbins), prof * bias


def evaluate_scheme(thresholds, bp, ref_ducy, nbins, ntrials, snr_final, ducy_max, wtsp, seed):
    """thresholding.py:1062-1110 with gen_next_using_thresh (:379-428) inlined."""
    nstages = len(bp)
    bias = snr_final / np.sqrt(nstages + 1)
    p0, p1 = _profiles(nbins, ref_ducy, bias)
    h0, var = simulate_folds(None, 0.0, p0, seed, 0, 1.0, ntrials)
    h1, _ = simulate_folds(None, 0.0, p1, seed, 1, 1.0, ntrials)
    prev, states = initial_state(), []
    for s in range(nstages):
        if s > 0 and (len(h0) == 0 or len(h1) == 0):
            break
        f0, v = simulate_folds(h0, var, p0, seed, 2 + 2 * s, 1.0, ntrials)
        g0 = prune_folds(f0, v, float(thresholds[s]), ducy_max, wtsp)
        f1, v = simulate_folds(h1, var, p1, seed, 3 + 2 * s, 1.0, ntrials)
        g1 = prune_folds(f1, v, float(thresholds[s]), ducy_max, wtsp)
        cur = get_next_state(prev, thresholds[s], len(g0) / len(f0), len(g1) / len(f1), bp[s])
        states.append(cur)
        prev, h0, h1, var = cur, g0, g1, v
    return np.array(states, dtype=STATE_DTYPE)


def determine_scheme(probs, bp, ref_ducy, nbins, ntrials, snr_final, ducy_max, wtsp, seed):
    """thresholding.py:1010-1059 with gen_next_using_surv_prob (:431-480) inlined; ties in
    the H0 ranking keep reverse row order (the product's rule)."""
    nstages = len(bp)
    bias = snr_final / np.sqrt(nstages + 1)
    p0, p1 = _profiles(nbins, ref_ducy, bias)
    h0, var = simulate_folds(None, 0.0, p0, seed, 0, 1.0, ntrials)
    h1, _ = simulate_folds(None, 0.0, p1, seed, 1, 1.0, ntrials)
    prev, states = initial_state(), []
    for s in range(nstages):
        if s > 0 and (len(h0) == 0 or len(h1) == 0):
            break
        f0, v = simulate_folds(h0, var, p0, seed, 2 + 2 * s, 1.0, ntrials)
        sc = fold_max_scores(f0, v, ducy_max, wtsp)
        n_surv = int(probs[s] * len(sc))
        order = np.argsort(sc, kind="stable")[::-1][:n_surv]
        thr = float(sc[order[-1]])
        g0 = np.ascontiguousarray(f0[order])
        f1, v = simulate_folds(h1, var, p1, seed, 3 + 2 * s, 1.0, ntrials)
        g1 = prune_folds(f1, v, thr, ducy_max, wtsp)
        cur = get_next_state(prev, thr, len(g0) / len(f0), len(g1) / len(f1), bp[s])
        states.append(cur)
        prev, h0, h1, var = cur, g0, g1, v
    return np.array(states, dtype=STATE_DTYPE)


def neighbouring_indices(beam, target, num):
    pos = int(np.searchsorted(beam, target))
    left = max(0, int(pos - num // 2))
    right = min(len(beam), left + num)
    left = max(0, right - num)
    return beam[left:right]


def dynamic_scheme(bp, ref_ducy, nbins, ntrials, nprobs, prob_min, snr_final, nthresholds,
                   ducy_max, wtsp, beam_width, mode, seed, thres_neigh, guess_path):
    """DynamicThresholdScheme (thresholding.py:702-962): __init__ / _init_legacy /
    _init_improved and run(), with the reference's loops (run_stage_legacy :588-648,
    pre_simulate_stage_folds :528-585, run_stage_improved :651-699). Returns the state table.
    ``guess_path`` is min(bound_scheme, trials_scheme), computed by the caller."""
    nstages = len(bp)
    bias = snr_final / np.sqrt(nstages + 1)
    p0, p1 = _profiles(nbins, ref_ducy, bias)
    thresholds = np.linspace(0.1, snr_final, nthresholds)
    probs = np.logspace(np.log10(prob_min), 0, nprobs)
    empty = np.ones(1, dtype=STATE_DTYPE)
    empty["threshold"] = -1.0
    empty["threshold_prev"] = -1.0
    states = np.full((nstages, nthresholds, nprobs), empty)

    def beam(istage):
        g = guess_path[istage]
        lo, hi = max(0, g - beam_width), min(snr_final, g + beam_width)
        return np.where((thresholds >= lo) & (thresholds <= hi))[0]

    folds_in: dict = {}
    h0, _ = simulate_folds(None, 0.0, p0, seed, 0, 1.0, ntrials)
    h1, _ = simulate_folds(None, 0.0, p1, seed, 1, 1.0, ntrials)
    var0 = 1.0
    init = initial_state()
    if mode == "improved":
        f0, v = simulate_folds(h0, var0, p0, seed, 2, 1.0, ntrials)
        f1, _ = simulate_folds(h1, var0, p1, seed, 3, 1.0, ntrials)
        sc0, sc1 = fold_max_scores(f0, v, ducy_max, wtsp), fold_max_scores(f1, v, ducy_max, wtsp)
        for i in beam(0):
            thr = thresholds[i]
            k0, k1 = np.nonzero(sc0 > thr)[0], np.nonzero(sc1 > thr)[0]
            cur = get_next_state(init, thr, len(k0) / len(f0), len(k1) / len(f1), bp[0])
            ip = int(np.digitize(cur["success_h1_cumul"], probs) - 1)
            if ip < 0:
                continue
            states[0, i, ip] = cur
            folds_in[(i, ip)] = (f0[k0], f1[k1], v)
    else:
        for i in beam(0):
            thr = thresholds[i]
            f0, v = simulate_folds(h0, var0, p0, seed, 2 + 2 * i, 1.0, ntrials)
            g0 = prune_folds(f0, v, thr, ducy_max, wtsp)
            f1, v = simulate_folds(h1, var0, p1, seed, 3 + 2 * i, 1.0, ntrials)
            g1 = prune_folds(f1, v, thr, ducy_max, wtsp)
            cur = get_next_state(init, thr, len(g0) / len(f0), len(g1) / len(f1), bp[0])
            ip = int(np.digitize(cur["success_h1_cumul"], probs) - 1)
            if ip < 0:
                continue
            states[0, i, ip] = cur
            folds_in[(i, ip)] = (g0, g1, v)
    for s in range(1, nstages):
        bcur, bprev = beam(s), beam(s - 1)
        sims = {}
        if mode == "improved":
            for j in bprev:
                for k in range(nprobs):
                    if states[s - 1, j, k]["is_empty"]:
                        continue
                    fs = folds_in.get((j, k))
                    if fs is None or len(fs[0]) == 0 or len(fs[1]) == 0:
                        continue
                    st = 4 + 2 * (((s - 1) * nthresholds + j) * nprobs + k)
                    f0, v = simulate_folds(fs[0], fs[2], p0, seed, st, 1.0, ntrials)
                    f1, _ = simulate_folds(fs[1], fs[2], p1, seed, st + 1, 1.0, ntrials)
                    sims[(j, k)] = (f0, f1, fold_max_scores(f0, v, ducy_max, wtsp),
                                    fold_max_scores(f1, v, ducy_max, wtsp), v)
        folds_out: dict = {}
        for i in bcur:
            for j in neighbouring_indices(bprev, i, thres_neigh):
                for k in range(nprobs):
                    prev = states[s - 1, j, k]
                    if prev["is_empty"]:
                        continue
                    fs = folds_in.get((j, k))
                    if fs is None or len(fs[0]) == 0 or len(fs[1]) == 0:
                        continue
                    thr = thresholds[i]
                    if mode == "improved":
                        f0, f1, sc0, sc1, v = sims[(j, k)]
                        k0, k1 = np.nonzero(sc0 > thr)[0], np.nonzero(sc1 > thr)[0]
                        g0, g1 = f0[k0], f1[k1]
                    else:
                        st = 2 + 2 * nthresholds + 2 * ((((s - 1) * nthresholds + i) * nthresholds
                                                         + j) * nprobs + k)
                        f0, v = simulate_folds(fs[0], fs[2], p0, seed, st, 1.0, ntrials)
                        g0 = prune_folds(f0, v, thr, ducy_max, wtsp)
                        f1, v = simulate_folds(fs[1], fs[2], p1, seed, st + 1, 1.0, ntrials)
                        g1 = prune_folds(f1, v, thr, ducy_max, wtsp)
                    cur = get_next_state(prev, thr, len(g0) / len(f0), len(g1) / len(f1), bp[s])
                    ip = int(np.digitize(cur["success_h1_cumul"], probs) - 1)
                    if ip < 0 or ip >= nprobs:
                        continue
                    ex = states[s, i, ip]
                    if ex["is_empty"] or cur["complexity_cumul"] < ex["complexity_cumul"]:
                        states[s, i, ip] = cur
                        folds_out[(i, ip)] = (g0, g1, v)
        folds_in = folds_out
    return states
