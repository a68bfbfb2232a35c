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

Circular basis (``circ_*``, 5 parameters [crackle, snap, jerk, accel, vel]):

- ``circ_mask``: ``circular.py:13-79,127-166``; ``circ_validate``: ``circular.py:82-124,
  397-440``.
- ``circ_branch_batch``: ``circular.py:204-394`` (crackle held at the parent value and
  expanded only for children in the snap-accel "hole" whose crackle still needs a split;
  kept children first, then each expanded child's crackle children).
- ``shift_taylor_circular``: ``transforms.py:248-313`` (values; ``:384-432`` keeps the
  errors unchanged); ``circ_resolve_batch``: ``circular.py:443-528`` with its in-place
  basis-flag writes (snap 1, crackle 2, crackle reset to 0 when the batch holds a Taylor
  cell); ``circ_transform_batch``: ``circular.py:669-705``; ``circ_ascend``:
  ``dyn_circular_taylor.py:636-678`` with ``circular.py:605-666``.
- One deliberate difference: the reference's circular branch leaves the d0 error of its
  output uninitialised (``np.empty``, ``circular.py:339``); it is never read by the search,
  and here it is 0.

Parity unpinned here: the reference runs these functions under numba ``fastmath``, so
its last-ulp choices (contraction, reassociation) are not reproduced; NumPy's cos/sin
(glibc) stand for the reference's.

Pinning: tests/test_circ_oracle.py checks the circular propagator against an analytic
orbit, the region masks and validation rules on hand-made cells, the branch against the
Taylor branch (no crackle split) and the crackle expansion's order and tiling, resolve
against the Taylor resolve (Taylor cells) and the analytic orbit, the flag writes, and a
level's survivors, flags, backtracks and transformed parameters against a child-by-child
brute force. tests/test_prune_oracle.py checks branching against the exact-tiling invariant
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

    def _fixed(self, level: int) -> tuple[float, float]:
        """snail.py:297-303: the span from the reference segment to the far edge of the
        segments so far, less the initial half-width, referenced to t_init."""
        t0_init, scale_init = self.get_coord(0)
        t0_next, scale_next = self.get_coord(level)
        left, right = t0_next - scale_next, t0_next + scale_next
        return t0_init, max(abs(left - t0_init), abs(right - t0_init)) - scale_init

    def get_current_coord(self, level: int, moving: bool = True) -> tuple[float, float]:
        """snail.py:286-303."""
        if level == 0:
            return self.get_coord(0)
        if moving:
            return self.get_coord(level - 1)[0], self.get_coord(level)[1]
        return self._fixed(level)

    def get_previous_coord(self, level: int, moving: bool = True) -> tuple[float, float]:
        """snail.py:307-326."""
        lv = level - 1
        if lv == 0 or moving:
            return self.get_coord(lv)
        return self._fixed(lv)

    def get_report_coord(self, moving: bool = True) -> tuple[float, float]:
        """snail.py:330-345: the fixed one keeps t0 of the last level."""
        lv = self.nsegments - 1
        if moving:
            return self.get_coord(lv)
        return self.get_coord(lv)[0], self._fixed(lv)[1]

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
    scale = coord_cur[1]
    P = poly_order
    pcur = leaves[:, :-2, 0]
    dcur = leaves[:, :-2, 1]
    f0 = leaves[:, -1, 0]
    dpf = step_f(P, scale, nbins, eta)
    dnew = dpf[None, :] * C_VAL / f0[:, None]
    fac = shift_factors(P, scale, nbins)
    shift_bins = np.abs(dcur - dnew) * (fac[None, :] * (f0 / C_VAL)[:, None])
    return odometer_children(pcur, dcur, dnew, shift_bins, eta, branch_max, leaves[:, -2, 0],
                             f0, leaves[:, -1, 1])


def odometer_children(pcur, dcur, dnew, shift_bins, eta: float, branch_max: int, d0, f0,
                      flag) -> tuple[np.ndarray, np.ndarray]:
    """``branch_param_padded`` per parameter (``psr_utils.py:347-379``), the keep rule for
    cells whose shift is below eta, and the odometer cartesian product per parent
    (``np_utils.py:174-238``). Children rows: branched parameters with their steps, then
    (d0, 0) and (f0, flag)."""
    n_batch, P = pcur.shape
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
        out[row:row + m, -2, 0] = d0[i]
        out[row:row + m, -1, 0] = f0[i]
        out[row:row + m, -1, 1] = flag[i]
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


def fixed_resolve_batch(leaves, coord_add, coord_init, counts_init, limits, nbins):
    """poly_taylor_fixed_resolve_batch (taylor.py:244-277): the leaf is referenced to
    t_init; velocity and delay at t_add are used whole."""
    t0_init, t0_add = coord_init[0], coord_add[0]
    vec = leaves[:, :-1, 0]
    f0 = leaves[:, -1, 0]
    d_add = shift_taylor(vec, t0_add - t0_init)
    accel = d_add[:, -3]
    freq = f0 * (1 - d_add[:, -2] / C_VAL)
    delay = d_add[:, -1] / C_VAL
    phase = phase_idx_vec(t0_add - t0_init, f0, nbins, delay)
    return nearest_2d(accel, freq, counts_init, limits), phase


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

    def __init__(self, size: int, P: int, B: int, fold_dtype=np.float32) -> None:
        self.size = size
        self.valid = 0
        self.leaves = np.empty((size, P + 2, 2))
        self.folds = np.empty((size, 2, B), dtype=fold_dtype)
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


def _fold_and_score(seg, phase, folds, parents, widths):
    """Children folds and scores: real rows (integer shift), or complex64 rows
    (fractional phase rotation, irfft before scoring) when the tree's folds are complex."""
    if np.iscomplexobj(folds):
        from oracle import complex_oracle as co

        comb = co.cshift_add_batch(seg, phase, folds, parents)
        return comb, co.cscore_batch(comb, widths)
    comb = shift_add_batch(seg, phase, folds, parents)
    return comb, orc.snr_score_batch(comb, widths)


def _ascend_sum(parts, like: np.ndarray) -> np.ndarray:
    """``shift_add_ascend_batch`` (common.py:432-459) or its complex form (:463-490): the
    segments' rows, each rotated by its phase, summed in order from zero."""
    if np.iscomplexobj(like):
        from oracle import complex_oracle as co

        return co.cascend_accumulate([r for r, _ in parts], [p for _, p in parts])
    B = like.shape[-1]
    res = np.zeros((len(like), 2, B), dtype=np.float32)
    for rows, phase in parts:
        s = shift_int_vec(phase, B)
        src = (np.arange(B)[None, :] - s[:, None]) % B
        rolled = np.take_along_axis(rows, np.broadcast_to(src[:, None, :], rows.shape), axis=2)
        res = (res + rolled).astype(np.float32)
    return res


def _score_folds(folds, widths):
    if np.iscomplexobj(folds):
        from oracle import complex_oracle as co

        return co.cscore_batch(folds, widths)
    return orc.snr_score_batch(folds, widths)


@dataclass
class LevelStats:
    n_leaves: int
    n_leaves_phy: int
    score_min: float
    score_max: float
    n_surv: int


def level_loop(tree: Tree, fold_segment: np.ndarray, threshold: float, cfg: dict,
               max_sugg: int, batch_size: int, branch, resolve, transform):
    """The batch loop of ``pruning_iteration_batched`` (``prune.py:165-332``) for a basis
    given by ``branch(leaves) -> (children, origins)``, ``resolve(children) -> (idx,
    phase)`` and ``transform(children) -> leaves at coord_next`` (validation is the
    identity for the polynomial bases). Returns (new tree, stats, survivor child ids)."""
    P = cfg["poly_order"]
    B = tree.folds.shape[-1]
    buf = TreeBuffer(max_sugg, P, B, tree.folds.dtype)
    n_leaves = 0
    smin, smax = math.inf, -math.inf
    cur_thr = threshold
    nb = tree.size
    bs = max(1, min(batch_size, nb))
    child_base = 0
    for b0 in range(0, nb, bs):
        b1 = min(b0 + bs, nb)
        ch, org = branch(tree.leaves[b0:b1])
        nl = len(ch)
        n_leaves += nl
        if nl == 0:
            continue
        idx, phase = resolve(ch)
        seg = seg_rows(fold_segment, idx)
        parents = np.arange(b0, b1)[org]
        comb, scores = _fold_and_score(seg, phase, tree.folds, parents, cfg["widths"])
        smin = min(smin, float(scores.min()))
        smax = max(smax, float(scores.max()))
        # numba compares the f32 scores with the f64 threshold in f64
        passing = scores.astype(np.float64) >= max(threshold, cur_thr)
        fi = np.where(passing)[0]
        if len(fi) == 0:
            child_base += nl
            continue
        trans = transform(ch[fi])
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


def prune_level(tree: Tree, fold_segment: np.ndarray, coords: dict, threshold: float,
                cfg: dict, max_sugg: int, batch_size: int = 1024, moving: bool = True):
    """One ``pruning_iteration_batched`` with the Taylor plugin; on the fixed grid the
    resolve is the fixed one and the transform the identity (dyn_poly_taylor.py:483-568)."""
    P = cfg["poly_order"]
    if moving:
        def resolve(ch):
            return resolve_batch(ch, coords["add"], coords["cur"], coords["init"],
                                 cfg["counts_init"], cfg["param_limits"], cfg["nbins"])

        def transform(ch):
            return transform_batch(ch, coords["next"], coords["cur"], cfg["tiling"])
    else:
        def resolve(ch):
            return fixed_resolve_batch(ch, coords["add"], coords["init"], cfg["counts_init"],
                                       cfg["param_limits"], cfg["nbins"])

        def transform(ch):
            return ch
    return level_loop(
        tree, fold_segment, threshold, cfg, max_sugg, batch_size,
        lambda lv: branch_batch(lv, coords["cur"], cfg["nbins"], cfg["eta"], P,
                                cfg["branch_max"]),
        resolve, transform)


def ascend(tree: Tree, dyp_fold: np.ndarray, idx_segments, coord_segments, coord_cur,
           cfg: dict) -> Tree:
    """Re-integrate every survivor over all segments so far and rescore (new tree)."""
    n = tree.size
    if n == 0:
        return tree
    vec = tree.leaves[:, :-1, 0]
    f0 = tree.leaves[:, -1, 0]
    parts = []
    for iseg, (t0_seg, _) in zip(idx_segments, coord_segments):
        d = shift_taylor(vec, t0_seg - coord_cur[0])
        accel = d[:, -3]
        freq = f0 * (1 - d[:, -2] / C_VAL)
        delay = d[:, -1] / C_VAL
        phase = phase_idx_vec(t0_seg - coord_cur[0], f0, cfg["nbins"], delay)
        idx = nearest_2d(accel, freq, cfg["counts_init"], cfg["param_limits"])
        parts.append((seg_rows(dyp_fold[iseg], idx), phase))
    res = _ascend_sum(parts, tree.folds)
    return Tree(tree.leaves.copy(), res, _score_folds(res, cfg["widths"]),
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
              batch_size: int = 1024, nlevels: int | None = None, moving: bool = True):
    """``Pruning.execute`` (real folds, moving or fixed grid): seed, levels, final ascend
    (at coord_end, prune.py:680-683), report (the fixed grid's report drops its re-centring,
    dyn_poly_taylor.py:676-690).

    ``dyp_fold``: [nseg, *counts, 2, B]. Returns dict with the final tree, report, per-level
    stats and per-level survivor child ids.
    """
    nseg = dyp_fold.shape[0]
    scheme = MiddleOut(nseg, ref_seg, tseg)
    leaves = taylor_seed(param_arr, dparams, cfg["poly_order"])
    B = dyp_fold.shape[-1]
    folds = dyp_fold[scheme.ref_idx].reshape(-1, 2, B)
    folds = folds.astype(np.complex64 if np.iscomplexobj(folds) else np.float32)
    scores = _score_folds(folds, cfg["widths"])
    tree = Tree(leaves, folds.copy(), scores, np.zeros((len(leaves), cfg["poly_order"] + 2),
                                                       dtype=np.int32))
    stats, survivors, trees = [], [], []
    nl = nseg - 1 if nlevels is None else nlevels
    level = 0
    for level in range(1, nl + 1):
        seg_idx = scheme.get_segment_idx(level)
        coords = {
            "init": scheme.get_coord(0),
            "cur": scheme.get_current_coord(level, moving),
            "prev": scheme.get_previous_coord(level, moving),
            "next": scheme.get_coord(level),
            "add": scheme.get_segment_coord(level),
        }
        tree, st, surv = prune_level(tree, dyp_fold[seg_idx], coords, float(thresholds[level - 1]),
                                     cfg, max_sugg, batch_size, moving)
        stats.append(st)
        survivors.append(surv)
        trees.append(tree)
        if tree.size == 0:
            break
    final_level = level
    if tree.size > 0:
        idx_seg, coord_seg = scheme.get_segment_coords_so_far(final_level)
        tree = ascend(tree, dyp_fold, idx_seg, coord_seg,
                      scheme.get_previous_coord(final_level + 1, moving), cfg)
    return {"tree": tree, "report": report_batch(tree.leaves) if tree.size else tree.leaves,
            "stats": stats, "survivors": survivors, "trees": trees, "scheme": scheme}


# ============================================================================ circular basis
# Restatement of core/circular.py (branch :204-394, validate :82-124,397-440, resolve
# :443-528, transform :669-705, ascend :605-666) and transforms.py:248-313,384-432.
# Leaves have 7 rows: [crackle, snap, jerk, accel, vel, d0, (f0, flag)].

def circ_mask(leaves: np.ndarray, sig: float):
    """get_circ_taylor_mask / get_circ_mask (circular.py:13-79,127-166): boolean masks
    (snap-circular, crackle-circular, taylor)."""
    crackle = leaves[:, 0, 0]
    snap, dsnap = leaves[:, 1, 0], leaves[:, 1, 1]
    jerk, djerk = leaves[:, 2, 0], leaves[:, 2, 1]
    accel, daccel = leaves[:, 3, 0], leaves[:, 3, 1]
    is_sig_snap = np.abs(snap) > (sig * dsnap)
    is_sig_accel = np.abs(accel) > (sig * daccel)
    is_sig_jerk = np.abs(jerk) > (sig * djerk)
    m_snap = is_sig_snap & is_sig_accel & ((-snap * accel) > 0)
    m_crackle = ~m_snap & is_sig_jerk & ((-crackle * jerk) > 0)
    m_taylor = ~(m_snap | m_crackle)
    return m_snap, m_crackle, m_taylor


def circ_branch_batch(leaves, coord_cur, nbins, eta, branch_max, sig):
    """circ_taylor_branch_batch: crackle split deferred to hole leaves; within the batch the
    kept leaves come first (stage-1 order), then every expanded leaf's crackle children."""
    P = 5
    n = leaves.shape[0]
    scale = coord_cur[1]
    pcur = leaves[:, :-2, 0]
    dcur = leaves[:, :-2, 1]
    f0 = leaves[:, -1, 0]
    dpf = step_f(P, scale, nbins, eta)
    dnew = dpf[None, :] * C_VAL / f0[:, None]
    fac = shift_factors(P, scale, nbins)
    shift_bins = np.abs(dcur - dnew) * (fac[None, :] * (f0 / C_VAL)[:, None])
    counts = np.ones((n, P), dtype=np.int64)
    dact = np.empty((n, P))
    firsts = np.empty((n, P))
    for i in range(n):
        for j in range(1, P):
            dc, dn = dcur[i, j], dnew[i, j]
            if dc <= 0.0 or dn <= 0.0:
                raise ValueError("Both dparam_cur and dparam_new must be positive.")
            num = max(1, math.ceil(dc / dn - FLOAT_EPSILON))
            if num > branch_max:
                raise ValueError("Invalid input: increase branch_max.")
            counts[i, j] = num
            dact[i, j] = dc / num
            firsts[i, j] = (pcur[i, j] - dc / 2.0) + dact[i, j] / 2.0
    cr_num = np.empty(n, dtype=np.int64)
    for i in range(n):  # branch_dparam_crackle (psr_utils.py:382-398)
        dc, dn = dcur[i, 0], dnew[i, 0]
        if dc <= 0.0 or dn <= 0.0:
            raise ValueError("Both dparam_cur and dparam_new must be positive.")
        num = max(1, math.ceil(dc / dn - FLOAT_EPSILON))
        if num > branch_max:
            raise ValueError("Invalid input: increase branch_max.")
        cr_num[i] = num
        dact[i, 0] = dc / num
        firsts[i, 0] = pcur[i, 0]
    needs = shift_bins >= (eta - FLOAT_EPSILON)
    keep_par = ~needs
    counts[keep_par] = 1
    dact[keep_par] = dcur[keep_par]
    # stage 1 (crackle held at the parent value)
    rows, orgs = [], []
    for i in range(n):
        grids = [np.array([pcur[i, 0]])]
        for j in range(1, P):
            if keep_par[i, j]:
                grids.append(np.array([pcur[i, j]]))
            else:
                grids.append(firsts[i, j] + np.arange(counts[i, j]) * dact[i, j])
        mesh = np.meshgrid(*grids, indexing="ij")
        cart = np.stack([m.reshape(-1) for m in mesh], axis=1)
        rows.append(cart)
        orgs.append(np.full(len(cart), i, dtype=np.int64))
    cart = np.concatenate(rows)
    org = np.concatenate(orgs)
    dp = dact[org]
    # get_circ_taylor_mask_branch (circular.py:169-201)
    snap, jerk, accel = cart[:, 1], cart[:, 2], cart[:, 3]
    dsnap, djerk, daccel = dp[:, 1], dp[:, 2], dp[:, 3]
    m_snap = (np.abs(snap) > sig * dsnap) & (np.abs(accel) > sig * daccel) & ((-snap * accel) > 0)
    hole = ~m_snap & (np.abs(jerk) > sig * djerk)
    expand = hole & needs[org, 0]
    out_rows, out_org = [], []
    for s in np.where(~expand)[0]:
        lv = np.zeros((P + 2, 2))
        lv[:P, 0] = cart[s]
        lv[:P, 1] = dp[s]
        lv[P, 0] = leaves[org[s], -2, 0]
        lv[P + 1] = leaves[org[s], -1]
        out_rows.append(lv)
        out_org.append(org[s])
    for s in np.where(expand)[0]:
        i = org[s]
        dc, dn = dcur[i, 0], dnew[i, 0]
        num = max(1, math.ceil(dc / dn - FLOAT_EPSILON))
        if num > branch_max:
            raise ValueError("Invalid input: increase branch_max.")
        da = dc / num
        first = (cart[s, 0] - dc / 2.0) + da / 2.0
        for c in range(num):
            lv = np.zeros((P + 2, 2))
            lv[:P, 0] = cart[s]
            lv[:P, 1] = dp[s]
            lv[0, 0] = first + c * da
            lv[0, 1] = da
            lv[P, 0] = leaves[i, -2, 0]
            lv[P + 1] = leaves[i, -1]
            out_rows.append(lv)
            out_org.append(i)
    if not out_rows:
        return np.zeros((0, P + 2, 2)), np.zeros(0, dtype=np.int64)
    return np.stack(out_rows), np.array(out_org, dtype=np.int64)


def circ_validate(leaves, origins, p_orb_min, x_mass_const, val_sig):
    """circ_taylor_validate_batch / circ_validate_batch: indices of kept leaves."""
    snap, dsnap = leaves[:, 1, 0], leaves[:, 1, 1]
    accel, daccel = leaves[:, 3, 0], leaves[:, 3, 1]
    omega_max_sq = (2 * np.pi / p_orb_min) ** 2
    sig_s = np.abs(snap) > val_sig * dsnap
    sig_a = np.abs(accel) > val_sig * daccel
    phys = (-snap * accel) > 0.0
    keep = np.ones(len(snap), dtype=bool)
    keep[sig_s & sig_a & ~phys] = False
    region = sig_s & sig_a & phys
    if np.any(region):
        idx = np.flatnonzero(region)
        om = -snap[idx] / accel[idx]
        lim = x_mass_const * (om ** (2.0 / 3.0))
        keep[idx] = (om <= omega_max_sq) & (np.abs(accel[idx]) <= lim)
    k = np.flatnonzero(keep)
    return leaves[k], origins[k], k


def shift_taylor_circular(vec: np.ndarray, dt: float, in_hole: bool) -> np.ndarray:
    """transforms.py:248-313 (vec rows [d5, d4, d3, d2, d1, d0])."""
    d5, d4, d3, d2, d1, d0 = (vec[:, k] for k in range(6))
    osq = -d5 / d3 if in_hole else -d4 / d2
    om = np.sqrt(osq)
    odt = om * dt
    c = np.cos(odt)
    s = np.sin(odt)
    d2j = d2 * c + (d3 / om) * s
    d3j = d3 * c - (d2 * om) * s
    d4j = -osq * d2j
    d5j = -osq * d3j
    d1ci = -d3 / osq
    d1cj = -d3j / osq
    d1j = d1cj + (d1 - d1ci)
    d0cj = -d2j / osq
    d0ci = -d2 / osq
    d0j = d0cj + (d0 - d0ci) + (d1 - d1ci) * dt
    return np.stack([d5j, d4j, d3j, d2j, d1j, d0j], axis=1)


def circ_shift_all(vec, dt, m_snap, m_crackle, m_taylor):
    out = np.empty_like(vec)
    if m_snap.any():
        out[m_snap] = shift_taylor_circular(vec[m_snap], dt, False)
    if m_crackle.any():
        out[m_crackle] = shift_taylor_circular(vec[m_crackle], dt, True)
    if m_taylor.any():
        out[m_taylor] = shift_taylor(vec[m_taylor], dt)
    return out


def circ_resolve_values(leaves, coord_add, coord_cur, coord_init, sig):
    """The (accel, freq, delay) a leaf resolves to at the added segment, and the masks
    (circ_taylor_resolve_batch, circular.py:443-505)."""
    t0_cur, t0_init, t0_add = coord_cur[0], coord_init[0], coord_add[0]
    vec = leaves[:, :-1, 0]
    f0 = leaves[:, -1, 0]
    ms, mc, mt = circ_mask(leaves, sig)
    d_add = circ_shift_all(vec, t0_add - t0_cur, ms, mc, mt)
    d_init = circ_shift_all(vec, t0_init - t0_cur, ms, mc, mt)
    accel = d_add[:, -3]
    vel = d_add[:, -2] - d_init[:, -2]
    freq = f0 * (1 - vel / C_VAL)
    delay = (d_add[:, -1] - d_init[:, -1]) / C_VAL
    return accel, freq, delay, (ms, mc, mt)


def circ_resolve_batch(leaves, coord_add, coord_cur, coord_init, counts_init, limits, nbins,
                       sig):
    """circ_taylor_resolve_batch (circular.py:443-528), including its in-place basis-flag
    writes: snap -> 1, crackle -> 2, and crackle leaves reset to 0 when the batch also holds
    a Taylor leaf (the reference indexes idx_circ_crackle in the Taylor branch, :510)."""
    t0_init, t0_add = coord_init[0], coord_add[0]
    f0 = leaves[:, -1, 0]
    accel, freq, delay, (ms, mc, mt) = circ_resolve_values(leaves, coord_add, coord_cur,
                                                           coord_init, sig)
    if ms.any():
        leaves[ms, -1, 1] = 1
    if mc.any():
        leaves[mc, -1, 1] = 2
    if mt.any():
        leaves[mc, -1, 1] = 0
    phase = phase_idx_vec(t0_add - t0_init, f0, nbins, delay)
    return nearest_2d(accel, freq, counts_init, limits), phase


def circ_fixed_resolve_batch(leaves, coord_add, coord_init, counts_init, limits, nbins, sig):
    """circ_taylor_fixed_resolve_batch (circular.py:531-590): shifts from t_init, velocity
    and delay used whole, and no basis-flag writes."""
    t0_init, t0_add = coord_init[0], coord_add[0]
    vec = leaves[:, :-1, 0]
    f0 = leaves[:, -1, 0]
    ms, mc, mt = circ_mask(leaves, sig)
    d_add = circ_shift_all(vec, t0_add - t0_init, ms, mc, mt)
    accel = d_add[:, -3]
    freq = f0 * (1 - d_add[:, -2] / C_VAL)
    delay = d_add[:, -1] / C_VAL
    phase = phase_idx_vec(t0_add - t0_init, f0, nbins, delay)
    return nearest_2d(accel, freq, counts_init, limits), phase


def circ_transform_batch(leaves, coord_next, coord_cur, tiling, sig):
    ms, mc, mt = circ_mask(leaves, sig)
    dt = coord_next[0] - coord_cur[0]
    out = leaves.copy()
    vals = leaves[:, :-1, 0]
    if ms.any():
        out[ms, :-1, 0] = shift_taylor_circular(vals[ms], dt, False)
    if mc.any():
        out[mc, :-1, 0] = shift_taylor_circular(vals[mc], dt, True)
    if mt.any():
        out[mt] = transform_batch(leaves[mt], coord_next, coord_cur, tiling)
    return out


def circ_prune_level(tree: Tree, fold_segment, coords, threshold, cfg, max_sugg,
                     batch_size=1024, moving=True):
    """pruning_iteration_batched with the circular plugin (prune.py:165-332)."""
    P = 5
    B = tree.folds.shape[-1]
    buf = TreeBuffer(max_sugg, P, B, tree.folds.dtype)
    n_leaves = n_phy = 0
    smin, smax = math.inf, -math.inf
    cur_thr = threshold
    nb = tree.size
    bs = max(1, min(batch_size, nb))
    child_base = 0
    for b0 in range(0, nb, bs):
        b1 = min(b0 + bs, nb)
        ch, org = circ_branch_batch(tree.leaves[b0:b1], coords["cur"], cfg["nbins"], cfg["eta"],
                                    cfg["branch_max"], cfg["prop_sig"])
        n_leaves += len(ch)
        if len(ch) == 0:
            continue
        ch, org, _ = circ_validate(ch, org, cfg["p_orb_min"], cfg["x_mass_const"],
                                   cfg["val_sig"])
        n_phy += len(ch)
        if len(ch) == 0:
            continue
        if moving:
            idx, phase = circ_resolve_batch(ch, coords["add"], coords["cur"], coords["init"],
                                            cfg["counts_init"], cfg["param_limits"],
                                            cfg["nbins"], cfg["prop_sig"])
        else:
            idx, phase = circ_fixed_resolve_batch(ch, coords["add"], coords["init"],
                                                  cfg["counts_init"], cfg["param_limits"],
                                                  cfg["nbins"], cfg["prop_sig"])
        seg = seg_rows(fold_segment, idx)
        parents = np.arange(b0, b1)[org]
        comb, scores = _fold_and_score(seg, phase, tree.folds, parents, cfg["widths"])
        smin = min(smin, float(scores.min()))
        smax = max(smax, float(scores.max()))
        fi = np.where(scores.astype(np.float64) >= max(threshold, cur_thr))[0]
        if len(fi):
            trans = (circ_transform_batch(ch[fi], coords["next"], coords["cur"], cfg["tiling"],
                                          cfg["prop_sig"]) if moving else ch[fi])
            bt = np.empty((len(fi), P + 2), dtype=np.int32)
            bt[:, 0] = parents[fi]
            bt[:, 1: idx.shape[1] + 1] = idx[fi]
            bt[:, -1] = phase[fi].astype(np.int32)
            cur_thr = buf.add_batch(trans, comb[fi], scores[fi], bt, child_base + fi, cur_thr)
        child_base += len(ch)
    v = buf.valid
    new = Tree(buf.leaves[:v].copy(), buf.folds[:v].copy(), buf.scores[:v].copy(),
               buf.backtracks[:v].copy())
    stats = LevelStats(n_leaves, n_phy, smin if math.isfinite(smin) else 0.0,
                       smax if math.isfinite(smax) else 0.0, v)
    return new, stats, buf.origin[:v].copy()


def circ_ascend(tree: Tree, dyp_fold, idx_segments, coord_segments, coord_cur, cfg) -> Tree:
    n = tree.size
    if n == 0:
        return tree
    vec = tree.leaves[:, :-1, 0]
    f0 = tree.leaves[:, -1, 0]
    ms, mc, mt = circ_mask(tree.leaves, cfg["prop_sig"])
    parts = []
    for iseg, (t0_seg, _) in zip(idx_segments, coord_segments):
        d = circ_shift_all(vec, t0_seg - coord_cur[0], ms, mc, mt)
        accel = d[:, -3]
        freq = f0 * (1 - d[:, -2] / C_VAL)
        delay = d[:, -1] / C_VAL
        phase = phase_idx_vec(t0_seg - coord_cur[0], f0, cfg["nbins"], delay)
        idx = nearest_2d(accel, freq, cfg["counts_init"], cfg["param_limits"])
        parts.append((seg_rows(dyp_fold[iseg], idx), phase))
    res = _ascend_sum(parts, tree.folds)
    return Tree(tree.leaves.copy(), res, _score_folds(res, cfg["widths"]),
                tree.backtracks.copy(), scores_ep=tree.scores.copy())


def circ_prune_run(dyp_fold, param_arr, dparams, tseg, thresholds, ref_seg, cfg, max_sugg,
                   batch_size=1024, nlevels=None, moving=True):
    """Pruning.execute with the circular plugin (real folds, moving or fixed grid)."""
    nseg = dyp_fold.shape[0]
    scheme = MiddleOut(nseg, ref_seg, tseg)
    leaves = taylor_seed(param_arr, dparams, 5)
    B = dyp_fold.shape[-1]
    folds = dyp_fold[scheme.ref_idx].reshape(-1, 2, B)
    folds = folds.astype(np.complex64 if np.iscomplexobj(folds) else np.float32)
    tree = Tree(leaves, folds.copy(), _score_folds(folds, cfg["widths"]),
                np.zeros((len(leaves), 7), dtype=np.int32))
    stats, survivors, trees = [], [], []
    nl = nseg - 1 if nlevels is None else nlevels
    level = 0
    for level in range(1, nl + 1):
        coords = {"init": scheme.get_coord(0), "cur": scheme.get_current_coord(level, moving),
                  "next": scheme.get_coord(level), "add": scheme.get_segment_coord(level)}
        tree, st, surv = circ_prune_level(tree, dyp_fold[scheme.get_segment_idx(level)], coords,
                                          float(thresholds[level - 1]), cfg, max_sugg,
                                          batch_size, moving)
        stats.append(st)
        survivors.append(surv)
        trees.append(tree)
        if tree.size == 0:
            break
    if tree.size > 0:
        idx_seg, coord_seg = scheme.get_segment_coords_so_far(level)
        tree = circ_ascend(tree, dyp_fold, idx_seg, coord_seg,
                           scheme.get_previous_coord(level + 1, moving), cfg)
    return {"tree": tree, "report": report_batch(tree.leaves) if tree.size else tree.leaves,
            "stats": stats, "survivors": survivors, "trees": trees, "scheme": scheme}
