"""CPU oracle of the EP (enumerate-and-prune) Taylor search -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this.
It restates, in NumPy with element-wise IEEE operations (no BLAS, no contraction), the
reference's real-fold Taylor pruning with a moving grid:

- ``MiddleOut``: ``utils/snail.py:196-343`` (segment order, coordinates).
- ``taylor_seed``: ``core/taylor.py:14-65`` + ``core/common.py:14-33``.
- ``branch_batch``: ``core/taylor.py:68-169`` with ``psr_utils.py:131-143,227-248,347-379``
  and ``np_utils.py:174-238`` (odometer cartesian product, parent-major).
- ``resolve_batch``: ``core/taylor.py:172-241`` with ``psr_utils.py:517-554``.
- ``shift_add_batch``: ``core/common.py:375-400``; scores: ``scoring.py:249-282`` (C oracle).
- ``transform_batch``: ``core/taylor.py:323-339`` / ``transforms.py:93-137``.
- ``prune_level``: ``prune.py:165-332`` (batches of 1024 parents, threshold
  ``max(threshold, current_threshold)``, backtracks, ``WorldTree.add_batch`` with the
  overflow rule of ``world_tree.py:482-551``).
- ``ascend``: ``dyn_poly_taylor.py:587-628``, ``taylor.py:280-320``, ``common.py:432-459``.
- ``report_batch``: ``taylor.py:342-360``.

Pinning: tests/test_prune_oracle.py checks branching against the exact-tiling invariant
and closed-form child counts, resolve against exact rational arithmetic, the level loop's
survivor set against a leaf-by-leaf brute force (every child scored independently, kept
iff score >= threshold) when the buffer does not overflow, the overflow rule against a
direct sort-based restatement, and the whole search against the injected signal.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from oracle import oracle as orc

C_VAL = 299792458.0
FLOAT_EPSILON = 1e-6


def powi(a: float, b: int) -> float:
    return orc.powi(float(a), int(b))


def fact(n: int) -> float:
    return float(math.factorial(n))


# ----------------------------------------------------------------------------- scheme
class MiddleOut:
    """``MiddleOutScheme`` with stride 1 (``snail.py:196-343``)."""

    def __init__(self, nsegments: int, ref_idx: int, tsegment: float) -> None:
        if nsegments <= 0:
            msg = f"nsegments must be greater than 0, got {nsegments}."
            raise ValueError(msg)
        if ref_idx < 0 or ref_idx >= nsegments:
            msg = f"ref_idx must be in [0, {nsegments}), got {ref_idx}."
            raise ValueError(msg)
        self.nsegments = nsegments
        self.ref_idx = ref_idx
        self.tsegment = tsegment
        self.data = np.argsort(np.abs(np.arange(nsegments) - ref_idx), kind="mergesort")

    def get_segment_idx(self, level: int) -> int:
        return int(self.data[level])

    def get_coord(self, level: int) -> tuple[float, float]:
        till = self.data[: level + 1]
        ref = (int(till.min()) + int(till.max()) + 1) / 2
        scale = ref - int(till.min())
        return ref * self.tsegment, scale * self.tsegment

    def get_segment_coord(self, level: int) -> tuple[float, float]:
        return (self.get_segment_idx(level) + 0.5) * self.tsegment, 0.5 * self.tsegment

    def get_current_coord(self, level: int) -> tuple[float, float]:
        if level == 0:
            return self.get_coord(0)
        return self.get_coord(level - 1)[0], self.get_coord(level)[1]

    def get_previous_coord(self, level: int) -> tuple[float, float]:
        return self.get_coord(level - 1)

    def get_segment_coords_so_far(self, level: int):
        idx = np.array([self.get_segment_idx(i) for i in range(level + 1)], dtype=np.int64)
        res = np.zeros((level + 1, 2))
        for i in range(level + 1):
            res[i] = self.get_segment_coord(i)
        return idx, res


# ----------------------------------------------------------------------------- tree
@dataclass
class Tree:
    leaves: np.ndarray        # f64 [n, P+2, 2]
    folds: np.ndarray         # f32 [n, 2, B]
    scores: np.ndarray        # f32 [n]
    backtracks: np.ndarray    # i32 [n, P+2]
    scores_ep: np.ndarray = field(default=None)

    def __post_init__(self) -> None:
        if self.scores_ep is None:
            self.scores_ep = self.scores.copy()

    @property
    def size(self) -> int:
        return len(self.scores)


def taylor_seed(param_arr: list[np.ndarray], dparams: np.ndarray, poly_order: int) -> np.ndarray:
    """Seed leaves from the last FFA grid (cartesian product, first parameter slowest)."""
    mesh = np.meshgrid(*param_arr, indexing="ij")
    cart = np.stack([m.reshape(-1) for m in mesh], axis=1)
    n = cart.shape[0]
    leaves = np.zeros((n, poly_order + 2, 2), dtype=np.float64)
    nparams = len(param_arr)
    # leaves[:, :-3] = (d_P .. d_2) from the FFA params except freq (with their steps)
    for k in range(nparams - 1):
        leaves[:, k, 0] = cart[:, k]
        leaves[:, k, 1] = dparams[k]
    f0 = cart[:, -1]
    df = np.full(n, dparams[-1])
    leaves[:, -3, 0] = 0.0
    leaves[:, -3, 1] = df * (C_VAL / f0)
    leaves[:, -2, 0] = 0.0
    leaves[:, -1, 0] = f0
    leaves[:, -1, 1] = 0.0
    return leaves


def step_f(n_params: int, tobs: float, nbins: int, eta: float) -> np.ndarray:
    """``poly_taylor_step_f`` with t_ref = 0, use_cheby (highest derivative first)."""
    dphi = eta / nbins
    out = [(2**k) * (dphi * fact(k + 1) / powi(tobs, k + 1)) for k in range(n_params)]
    return np.array(out[::-1])


def shift_factors(n_params: int, tobs: float, nbins: int) -> np.ndarray:
    """``(tobs)^(k+1) * nbins / (k+1)! / 2^k`` for k = n-1..0 (``psr_utils.py:241-245``)."""
    out = []
    for k in range(n_params - 1, -1, -1):
        f = powi(tobs, k + 1) * nbins / fact(k + 1)
        out.append(f / (2**k))
    return np.array(out)


def branch_batch(leaves: np.ndarray, coord_cur, nbins: int, eta: float, poly_order: int,
                 branch_max: int) -> tuple[np.ndarray, np.ndarray]:
    n_batch = leaves.shape[0]
    scale = coord_cur[1]
    P = poly_order
    pcur = leaves[:, :-2, 0]
    dcur = leaves[:, :-2, 1]
    f0 = leaves[:, -1, 0]
    dpf = step_f(P, scale, nbins, eta)
    dnew = dpf[None, :] * C_VAL / f0[:, None]
    fac = shift_factors(P, scale, nbins)
    shift_bins = np.abs(dcur - dnew) * (fac[None, :] * (f0 / C_VAL)[:, None])
    counts = np.ones((n_batch, P), dtype=np.int64)
    dact = np.empty((n_batch, P))
    firsts = np.empty((n_batch, P))
    for i in range(n_batch):
        for j in range(P):
            dc, dn = dcur[i, j], dnew[i, j]
            if dc <= 0.0 or dn <= 0.0:
                msg = "Both dparam_cur and dparam_new must be positive."
                raise ValueError(msg)
            num = max(1, math.ceil(dc / dn - FLOAT_EPSILON))
            if num > branch_max:
                msg = "Invalid input: increase branch_max."
                raise ValueError(msg)
            da = dc / num
            counts[i, j] = num
            dact[i, j] = da
            firsts[i, j] = (pcur[i, j] - dc / 2.0) + da / 2.0
    keep = shift_bins < (eta - FLOAT_EPSILON)
    counts[keep] = 1
    dact[keep] = dcur[keep]
    nchild = counts.prod(axis=1)
    total = int(nchild.sum())
    out = np.zeros((total, P + 2, 2))
    origins = np.repeat(np.arange(n_batch), nchild)
    row = 0
    for i in range(n_batch):
        grids = []
        for j in range(P):
            if keep[i, j]:
                grids.append(np.array([pcur[i, j]]))
            else:
                grids.append(firsts[i, j] + np.arange(counts[i, j]) * dact[i, j])
        mesh = np.meshgrid(*grids, indexing="ij")
        cart = np.stack([m.reshape(-1) for m in mesh], axis=1)
        m = cart.shape[0]
        out[row:row + m, :-2, 0] = cart
        out[row:row + m, :-2, 1] = dact[i]
        out[row:row + m, -2, 0] = leaves[i, -2, 0]
        out[row:row + m, -1, 0] = leaves[i, -1, 0]
        out[row:row + m, -1, 1] = leaves[i, -1, 1]
        row += m
    return out, origins


def shift_taylor(vals: np.ndarray, dt: float) -> np.ndarray:
    """Rows ``[d_P..d_0]`` translated by dt: out[:, i] = sum_j vals[:, j] T[i, j]."""
    n = vals.shape[1]
    T = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1):
            T[i, j] = powi(dt, i - j) / fact(i - j)
    out = np.zeros_like(vals)
    for i in range(n):
        acc = np.zeros(vals.shape[0])
        for j in range(n):
            acc = acc + vals[:, j] * T[i, j]
        out[:, i] = acc
    return out


def nearest_2d(accel, freq, counts_init, limits):
    n = len(accel)
    P = limits.shape[0]
    out = np.zeros((n, P), dtype=np.int64)
    n_acc, n_freq = int(counts_init[-2]), int(counts_init[-1])
    amin, amax = limits[-2]
    fmin, fmax = limits[-1]
    sa = n_acc / (amax - amin) if amax > amin else 0.0
    sf = n_freq / (fmax - fmin) if fmax > fmin else 0.0
    if sa > 0.0:
        ia = ((accel - amin) * sa + FLOAT_EPSILON).astype(np.float64)
        ia = np.trunc(ia).astype(np.int64)
        out[:, -2] = np.clip(ia, 0, n_acc - 1)
    if sf > 0.0:
        fi = np.trunc((freq - fmin) * sf + FLOAT_EPSILON).astype(np.int64)
        out[:, -1] = np.clip(fi, 0, n_freq - 1)
    return out


def phase_idx_vec(t: float, f: np.ndarray, nbins: int, delay: np.ndarray) -> np.ndarray:
    phase = (t - delay) * f
    phase = phase - np.floor(phase)
    ip = phase * nbins
    ip[ip >= nbins] = 0.0
    return ip


def shift_int_vec(x: np.ndarray, nbins: int) -> np.ndarray:
    r = (x.astype(np.float32) + np.float32(0.5)).astype(np.int64)
    r[r == nbins] = 0
    return r


def resolve_batch(leaves, coord_add, coord_cur, coord_init, counts_init, limits, nbins):
    t0_cur, t0_init, t0_add = coord_cur[0], coord_init[0], coord_add[0]
    vec = leaves[:, :-1, 0]
    f0 = leaves[:, -1, 0]
    d_add = shift_taylor(vec, t0_add - t0_cur)
    d_init = shift_taylor(vec, t0_init - t0_cur)
    accel = d_add[:, -3]
    vel = d_add[:, -2] - d_init[:, -2]
    freq = f0 * (1 - vel / C_VAL)
    delay = (d_add[:, -1] - d_init[:, -1]) / C_VAL
    phase = phase_idx_vec(t0_add - t0_init, f0, nbins, delay)
    idx = nearest_2d(accel, freq, counts_init, limits)
    return idx, phase


def seg_rows(fold_segment: np.ndarray, idx: np.ndarray) -> np.ndarray:
    """``load_prune_folds_Nd``: index 0 on all but the (accel, freq) axes."""
    if fold_segment.ndim == 3:  # (nfreq, 2, B)
        return fold_segment[idx[:, -1]]
    fs = fold_segment.reshape(-1, *fold_segment.shape[-4:])[0]  # leading axes at index 0
    return fs[idx[:, -2], idx[:, -1]]


def shift_add_batch(seg: np.ndarray, shifts_f64: np.ndarray, folds: np.ndarray,
                    parents: np.ndarray) -> np.ndarray:
    B = seg.shape[-1]
    s = shift_int_vec(shifts_f64, B)
    j = np.arange(B)
    src = (j[None, :] - s[:, None]) % B
    rolled = np.take_along_axis(seg, np.broadcast_to(src[:, None, :], seg.shape), axis=2)
    return (folds[parents] + rolled).astype(np.float32)


def transform_batch(leaves, coord_next, coord_cur, tiling: str = "aggressive"):
    dt = coord_next[0] - coord_cur[0]
    out = np.zeros_like(leaves)
    vals = leaves[:, :-1, 0]
    errs = leaves[:, :-1, 1]
    out[:, :-1, 0] = shift_taylor(vals, dt)
    n = vals.shape[1]
    T = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1):
            T[i, j] = powi(dt, i - j) / fact(i - j)
    if tiling == "aggressive":
        out[:, :-1, 1] = errs * np.abs(np.diag(T))[None, :]
    elif tiling == "conservative":
        acc = np.zeros_like(errs)
        for i in range(n):
            a = np.zeros(errs.shape[0])
            for j in range(n):
                a = a + errs[:, j] * abs(T[i, j])
            acc[:, i] = a
        out[:, :-1, 1] = acc
    elif tiling == "quadrature":
        acc = np.zeros_like(errs)
        for i in range(n):
            a = np.zeros(errs.shape[0])
            for j in range(n):
                a = a + (errs[:, j] ** 2) * (T[i, j] ** 2)
            acc[:, i] = np.sqrt(a)
        out[:, :-1, 1] = acc
    else:
        msg = f"Invalid tiling strategy: {tiling}"
        raise ValueError(msg)
    out[:, -1] = leaves[:, -1]
    return out


class TreeBuffer:
    """``WorldTree`` of capacity ``size`` with ``add_batch`` / ``prune_on_overload``."""

    def __init__(self, size: int, P: int, B: int) -> None:
        self.size = size
        self.valid = 0
        self.leaves = np.empty((size, P + 2, 2))
        self.folds = np.empty((size, 2, B), dtype=np.float32)
        self.scores = np.empty(size, dtype=np.float32)
        self.backtracks = np.empty((size, P + 2), dtype=np.int32)
        self.origin = np.empty(size, dtype=np.int64)  # global child index (for tests)

    def _keep(self, mask: np.ndarray) -> None:
        idx = np.where(mask)[0]
        c = len(idx)
        for arr in (self.leaves, self.folds, self.scores, self.backtracks, self.origin):
            arr[:c] = arr[idx]
        self.valid = c

    def prune_on_overload(self, scores_batch: np.ndarray, cur_thr: float) -> float:
        n_add = len(scores_batch)
        total = self.valid + n_add
        scratch = np.concatenate([self.scores[: self.valid], scores_batch])
        k_idx = -self.size
        mid_idx = -(total // 2) - 1
        part = np.partition(scratch, [k_idx, mid_idx])
        topk = math.nextafter(float(part[k_idx]), math.inf)
        med = math.nextafter(float(part[mid_idx]), math.inf)
        eff = max(cur_thr, topk, med)
        self._keep(self.scores[: self.valid].astype(np.float64) >= eff)
        return eff

    def add_batch(self, leaves, folds, scores, bts, origin, cur_thr: float) -> float:
        n = len(scores)
        if n == 0:
            return cur_thr
        space = self.size - self.valid
        if n <= space:
            sl = slice(self.valid, self.valid + n)
            self.leaves[sl], self.folds[sl], self.scores[sl] = leaves, folds, scores
            self.backtracks[sl], self.origin[sl] = bts, origin
            self.valid += n
            return cur_thr
        eff = self.prune_on_overload(scores, cur_thr)
        pend = np.where(scores.astype(np.float64) >= eff)[0]
        space = self.size - self.valid
        m = min(len(pend), space)
        if m == 0:
            return eff
        pend = pend[:m]
        sl = slice(self.valid, self.valid + m)
        self.leaves[sl], self.folds[sl], self.scores[sl] = leaves[pend], folds[pend], scores[pend]
        self.backtracks[sl], self.origin[sl] = bts[pend], origin[pend]
        self.valid += m
        return eff


@dataclass
class LevelStats:
    n_leaves: int
    n_leaves_phy: int
    score_min: float
    score_max: float
    n_surv: int


def prune_level(tree: Tree, fold_segment: np.ndarray, coords: dict, threshold: float,
                cfg: dict, max_sugg: int, batch_size: int = 1024):
    """One ``pruning_iteration_batched``. Returns (new tree, stats, survivor child ids)."""
    P = cfg["poly_order"]
    B = tree.folds.shape[-1]
    buf = TreeBuffer(max_sugg, P, B)
    n_leaves = 0
    smin, smax = math.inf, -math.inf
    cur_thr = threshold
    nb = tree.size
    bs = max(1, min(batch_size, nb))
    child_base = 0
    for b0 in range(0, nb, bs):
        b1 = min(b0 + bs, nb)
        ch, org = branch_batch(tree.leaves[b0:b1], coords["cur"], cfg["nbins"], cfg["eta"], P,
                               cfg["branch_max"])
        nl = len(ch)
        n_leaves += nl
        if nl == 0:
            continue
        idx, phase = resolve_batch(ch, coords["add"], coords["cur"], coords["init"],
                                   cfg["counts_init"], cfg["param_limits"], cfg["nbins"])
        seg = seg_rows(fold_segment, idx)
        parents = np.arange(b0, b1)[org]
        comb = shift_add_batch(seg, phase, tree.folds, parents)
        scores = orc.snr_score_batch(comb, cfg["widths"])
        smin = min(smin, float(scores.min()))
        smax = max(smax, float(scores.max()))
        # numba compares the f32 scores with the f64 threshold in f64
        passing = scores.astype(np.float64) >= max(threshold, cur_thr)
        fi = np.where(passing)[0]
        if len(fi) == 0:
            child_base += nl
            continue
        trans = transform_batch(ch[fi], coords["next"], coords["cur"], cfg["tiling"])
        bt = np.empty((len(fi), P + 2), dtype=np.int32)
        bt[:, 0] = parents[fi]
        bt[:, 1: idx.shape[1] + 1] = idx[fi]
        bt[:, -1] = phase[fi].astype(np.int32)
        cur_thr = buf.add_batch(trans, comb[fi], scores[fi], bt, child_base + fi, cur_thr)
        child_base += nl
    v = buf.valid
    new = Tree(buf.leaves[:v].copy(), buf.folds[:v].copy(), buf.scores[:v].copy(),
               buf.backtracks[:v].copy())
    stats = LevelStats(n_leaves, n_leaves, smin if math.isfinite(smin) else 0.0,
                       smax if math.isfinite(smax) else 0.0, v)
    return new, stats, buf.origin[:v].copy()


def ascend(tree: Tree, dyp_fold: np.ndarray, idx_segments, coord_segments, coord_cur,
           cfg: dict) -> Tree:
    """Re-integrate every survivor over all segments so far and rescore (new tree)."""
    n = tree.size
    if n == 0:
        return tree
    vec = tree.leaves[:, :-1, 0]
    f0 = tree.leaves[:, -1, 0]
    B = tree.folds.shape[-1]
    res = np.zeros((n, 2, B), dtype=np.float32)
    for iseg, (t0_seg, _) in zip(idx_segments, coord_segments):
        d = shift_taylor(vec, t0_seg - coord_cur[0])
        accel = d[:, -3]
        freq = f0 * (1 - d[:, -2] / C_VAL)
        delay = d[:, -1] / C_VAL
        phase = phase_idx_vec(t0_seg - coord_cur[0], f0, cfg["nbins"], delay)
        idx = nearest_2d(accel, freq, cfg["counts_init"], cfg["param_limits"])
        rows = seg_rows(dyp_fold[iseg], idx)
        s = shift_int_vec(phase, B)
        src = (np.arange(B)[None, :] - s[:, None]) % B
        rolled = np.take_along_axis(rows, np.broadcast_to(src[:, None, :], rows.shape), axis=2)
        res = (res + rolled).astype(np.float32)
    return Tree(tree.leaves.copy(), res, orc.snr_score_batch(res, cfg["widths"]),
                tree.backtracks.copy(), scores_ep=tree.scores.copy())


def report_batch(leaves: np.ndarray) -> np.ndarray:
    out = leaves.copy()
    vals = leaves[:, :-3, 0]
    sigs = leaves[:, :-3, 1]
    v = leaves[:, -3, 0]
    dv = leaves[:, -3, 1]
    f0 = leaves[:, -1, 0]
    s = 1 - v / C_VAL
    out[:, :-3, 0] = vals / s[:, None]
    out[:, :-3, 1] = np.sqrt((sigs / s[:, None]) ** 2
                             + ((vals / (C_VAL * s[:, None] ** 2)) ** 2) * (dv[:, None] ** 2))
    out[:, -3, 0] = f0 * s
    out[:, -3, 1] = f0 * dv / C_VAL
    return out


def prune_run(dyp_fold: np.ndarray, param_arr: list[np.ndarray], dparams: np.ndarray,
              tseg: float, thresholds: np.ndarray, ref_seg: int, cfg: dict, max_sugg: int,
              batch_size: int = 1024, nlevels: int | None = None):
    """``Pruning.execute`` (moving grid, real folds): seed, levels, final ascend, report.

    ``dyp_fold``: [nseg, *counts, 2, B]. Returns dict with the final tree, report, per-level
    stats and per-level survivor child ids.
    """
    nseg = dyp_fold.shape[0]
    scheme = MiddleOut(nseg, ref_seg, tseg)
    leaves = taylor_seed(param_arr, dparams, cfg["poly_order"])
    B = dyp_fold.shape[-1]
    folds = dyp_fold[scheme.ref_idx].reshape(-1, 2, B).astype(np.float32)
    scores = orc.snr_score_batch(folds, cfg["widths"])
    tree = Tree(leaves, folds.copy(), scores, np.zeros((len(leaves), cfg["poly_order"] + 2),
                                                       dtype=np.int32))
    stats, survivors, trees = [], [], []
    nl = nseg - 1 if nlevels is None else nlevels
    level = 0
    for level in range(1, nl + 1):
        seg_idx = scheme.get_segment_idx(level)
        coords = {
            "init": scheme.get_coord(0),
            "cur": scheme.get_current_coord(level),
            "prev": scheme.get_previous_coord(level),
            "next": scheme.get_coord(level),
            "add": scheme.get_segment_coord(level),
        }
        tree, st, surv = prune_level(tree, dyp_fold[seg_idx], coords, float(thresholds[level - 1]),
                                     cfg, max_sugg, batch_size)
        stats.append(st)
        survivors.append(surv)
        trees.append(tree)
        if tree.size == 0:
            break
    final_level = level
    if tree.size > 0:
        idx_seg, coord_seg = scheme.get_segment_coords_so_far(final_level)
        tree = ascend(tree, dyp_fold, idx_seg, coord_seg, scheme.get_coord(final_level), cfg)
    return {"tree": tree, "report": report_batch(tree.leaves) if tree.size else tree.leaves,
            "stats": stats, "survivors": survivors, "trees": trees, "scheme": scheme}
