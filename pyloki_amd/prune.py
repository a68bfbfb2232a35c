"""EP (enumerate-and-prune) search on the GPU: ``Pruning`` and ``prune_dyp_tree``.

Mirrors ``pyloki.prune`` (``src/pyloki/prune.py``) for the Taylor and circular bases
(``poly_basis="taylor"``; circular when ``prune_poly_order == 5``), with real or complex
(``use_fourier=True``) folds and a moving or fixed grid (``use_moving_grid``): seed from the
FFA fold of the reference segment, one pruning level per remaining segment (middle-out
order), the final ascend (re-integration + rescoring) and the report transform.

On the fixed grid every leaf stays referenced to the reference segment: the coordinates
are ``MiddleOutScheme``'s fixed ones (``snail.py:286-345``), the resolve is the fixed
variant and the transform is the identity. Its report is the moving grid's report of
those leaves: the reference computes a re-centring there and drops the result
(``dyn_poly_taylor.py:676-690``, ``dyn_circular_taylor.py:725-740``); the Chebyshev report
converts with the fixed report coordinate (``dyn_poly_cheby.py:824-839``).

Every per-leaf step runs in the HIP library (``pyl_prune_*``). The host decides which
children survive: when the children passing the threshold fit the survivor buffer
(``max_sugg``) they are all kept (the reference's fast path); otherwise the reference's
batch-sequential overflow rule (``world_tree.py:482-551``) is replayed on the child scores
in its batch order, which selects exactly the reference's survivors. Survivors are always a
subsequence of the children in child order, so a boolean selection describes the new tree.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import torch

from pyloki_amd import _lib
from pyloki_amd.device import ptr, stream_ptr, to_device
from pyloki_amd.io.cands import (PruneResultWriter, PruneStats, PruneStatsCollection,
                                 mkdir_p)
from pyloki_amd.utils import transforms
from pyloki_amd.utils.misc import C_VAL
from pyloki_amd.utils.snail import MiddleOutScheme

INT64_MAX = 2**63 - 1
_TILING = {"aggressive": 0, "quadrature": 1, "conservative": 2}


class PruneParams(ctypes.Structure):
    _fields_ = [
        ("poly_order", ctypes.c_int32),
        ("nbins", ctypes.c_int32),
        ("branch_max", ctypes.c_int32),
        ("tiling", ctypes.c_int32),
        ("eta", ctypes.c_double),
        ("nparams_ffa", ctypes.c_int32),
        ("n_accel", ctypes.c_int32),
        ("n_freq", ctypes.c_int32),
        ("nw", ctypes.c_int32),
        ("accel_lim", ctypes.c_double * 2),
        ("freq_lim", ctypes.c_double * 2),
        ("widths", ctypes.c_int32 * 64),
        ("fixed_grid", ctypes.c_int32),
    ]


class PruneCoords(ctypes.Structure):
    _fields_ = [
        ("t0_init", ctypes.c_double),
        ("t0_cur", ctypes.c_double),
        ("scale_cur", ctypes.c_double),
        ("t0_next", ctypes.c_double),
        ("t0_add", ctypes.c_double),
        ("t0_prev", ctypes.c_double),
        ("scale_prev", ctypes.c_double),
        ("scale_next", ctypes.c_double),
    ]


class CircParams(ctypes.Structure):
    _fields_ = [
        ("prop_sig", ctypes.c_double),
        ("val_sig", ctypes.c_double),
        ("omega_max_sq", ctypes.c_double),
        ("x_mass_const", ctypes.c_double),
        ("batch_size", ctypes.c_int64),
    ]


class CircLevel(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in (
        "branch", "counts", "cr_first", "cr_num", "off1", "s1_parent", "s1_expand", "s1_pos",
        "c_s1", "c_sub", "c_keep", "c_region")] + [
        ("n_par", ctypes.c_int64), ("n_s1", ctypes.c_int64), ("n_child", ctypes.c_int64)]


REGION_TAYLOR = 2

# Sync-free level loop (pyl_ep_levels): flags, level record words and statuses.
EP_COMPLEX, EP_CHEBYSHEV = 1, 2
EP_SCAN_BLOCKS = 256
REC_NPAR, REC_NCHILD, REC_NSURV, REC_STATUS, REC_ERR, REC_HIST, REC_NVALID = range(7)
REC_WORDS = 12
ST_OK, ST_BRANCH, ST_CHILD_CAP, ST_OVERFLOW, ST_HIST_FULL, ST_BUDGET, ST_TREE_OVERFLOW = range(7)


class EpLoop(ctypes.Structure):
    _fields_ = [("cap_par", ctypes.c_int64), ("cap_child", ctypes.c_int64),
                ("cap_hist", ctypes.c_int64), ("max_sugg", ctypes.c_int64),
                ("leaves", ctypes.c_void_p * 2), ("folds", ctypes.c_void_p * 2),
                ("scores", ctypes.c_void_p * 2)] + [(name, ctypes.c_void_p) for name in (
        "hist", "leaves_cur", "branch", "counts", "nchild", "off", "c_parent", "c_row",
        "c_shift", "c_phase", "c_score", "surv", "part", "rec", "minmax", "timer", "budget")]
EP_PHASES = ("branch", "children", "select", "materialize", "budget")
REC_NKEPT = 10
BUDGET_SCORE, BUDGET_CNN = 1, 2
BUDGET_SEL_WORDS, BUDGET_HIST_BINS, BUDGET_STATS = 272, 1024, 1032


class EpBudget(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("nbins", ctypes.c_int32),
                ("complex_folds", ctypes.c_int32), ("nlevels", ctypes.c_int32),
                ("level0", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("budget", ctypes.c_int64)] + [
        (name, ctypes.c_void_p) for name in (
            "params", "keys", "feats", "kept", "bt_tmp", "sel", "stats", "shist", "part",
            "level_mask", "level_budget", "level_kind")]


EP_FOREST_MAX, EP_FOREST_TAB = 64, 6
TAB_PO, TAB_BO, TAB_CB, TAB_VB, TAB_SO, TAB_KO = range(6)


class EpForest(ctypes.Structure):
    _fields_ = [("ntree", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("tree_stride", ctypes.c_int64), ("max_sugg", ctypes.c_int64)] + [
        (name, ctypes.c_void_p) for name in ("tab", "minmax", "sel", "stats", "shist", "part",
                                             "kcount")]


class CircLoop(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in (
        "cr_first", "cr_num", "off1", "s1_parent", "s1_expand", "s1_pos", "held_off", "c_s1",
        "c_sub", "c_keep", "c_region", "valid", "batch_tot", "batch_nk", "batch_off",
        "has_taylor")]


class LoopBuffers:
    """Device buffers of the sync-free level loop, sized by capacities: two tree buffers
    of ``cap_par`` rows, the per-child arrays of one level (``cap_child``), the backtrack
    history (``cap_hist`` rows, appended level after level) and the level records."""

    _CHILD = (("c_parent", torch.int32), ("c_row", torch.int32), ("c_shift", torch.int32),
              ("c_phase", torch.float64), ("c_score", torch.float32))

    _CIRC_PAR = (("cr_first", torch.float64, 0), ("cr_num", torch.int32, 0),
                 ("off1", torch.int64, 1), ("batch_tot", torch.int64, 0),
                 ("batch_nk", torch.int64, 0), ("batch_off", torch.int64, 0),
                 ("has_taylor", torch.uint8, 0))
    _CIRC_CHILD = (("s1_parent", torch.int32, 0), ("s1_expand", torch.uint8, 0),
                   ("s1_pos", torch.int64, 0), ("held_off", torch.int64, 1),
                   ("c_s1", torch.int64, 0), ("c_sub", torch.int32, 0),
                   ("c_keep", torch.uint8, 0), ("c_region", torch.int8, 0),
                   ("valid", torch.int64, 0))

    def __init__(self, P: int, new_folds, cap_par: int, cap_child: int, cap_hist: int,
                 max_sugg: int, nlev: int, cheby: bool, circular: bool = False) -> None:
        dev = torch.device("cuda")
        f64, i32, i64 = torch.float64, torch.int32, torch.int64
        self.P = P
        self.leaves = [torch.empty((cap_par, P + 2, 2), dtype=f64, device=dev) for _ in (0, 1)]
        self.folds = [new_folds(cap_par, dev) for _ in (0, 1)]
        self.scores = [torch.empty(cap_par, dtype=torch.float32, device=dev) for _ in (0, 1)]
        self.leaves_cur = (torch.empty((cap_par, P + 2, 2), dtype=f64, device=dev) if cheby
                           else None)
        self.branch = torch.empty((cap_par, P, 2), dtype=f64, device=dev)
        self.counts = torch.empty((cap_par, P), dtype=i32, device=dev)
        self.nchild = torch.empty(cap_par, dtype=i64, device=dev)
        self.off = torch.empty(cap_par + 1, dtype=i64, device=dev)
        self.surv = torch.empty(max_sugg, dtype=i64, device=dev)
        self.part = torch.empty(EP_SCAN_BLOCKS, dtype=i64, device=dev)
        self.rec = torch.empty((nlev + 1, REC_WORDS), dtype=i64, device=dev)
        self.minmax = torch.empty((nlev, 2), dtype=torch.float32, device=dev)
        self.s = EpLoop()
        self.s.cap_par, self.s.max_sugg = cap_par, max_sugg
        self.c = CircLoop() if circular else None
        if circular:
            for name, dt, extra in self._CIRC_PAR:
                t = torch.empty(cap_par + extra, dtype=dt, device=dev)
                setattr(self, name, t)
                setattr(self.c, name, ptr(t))
        for j in (0, 1):
            self.s.leaves[j] = ptr(self.leaves[j])
            self.s.folds[j] = ptr(self.folds[j])
            self.s.scores[j] = ptr(self.scores[j])
        for name in ("leaves_cur", "branch", "counts", "nchild", "off", "surv", "part", "rec",
                     "minmax"):
            setattr(self.s, name, ptr(getattr(self, name)))
        self.set_children(cap_child)
        self.new_hist(cap_hist)

    @property
    def cap_par(self) -> int:
        return int(self.s.cap_par)

    @property
    def cap_child(self) -> int:
        return int(self.s.cap_child)

    @property
    def nlev(self) -> int:
        return int(self.minmax.shape[0])

    def set_children(self, cap_child: int) -> None:
        dev = self.nchild.device
        for name, dt in self._CHILD:
            t = torch.empty(cap_child, dtype=dt, device=dev)
            setattr(self, name, t)
            setattr(self.s, name, ptr(t))
        if self.c is not None:
            for name, dt, extra in self._CIRC_CHILD:
                t = torch.empty(cap_child + extra, dtype=dt, device=dev)
                setattr(self, name, t)
                setattr(self.c, name, ptr(t))
        self.s.cap_child = cap_child

    def new_hist(self, cap_hist: int) -> None:
        """A fresh history (views of the old one stay valid)."""
        self.hist = torch.empty((cap_hist, self.P + 2), dtype=torch.int32,
                                device=self.nchild.device)
        self.s.hist = ptr(self.hist)
        self.s.cap_hist = cap_hist
        self.hist_next = 0

    def holds(self, tree) -> int:
        """Index of the tree buffer that ``tree`` is a view of, else -1."""
        for j in (0, 1):
            if tree.leaves.numel() and tree.leaves.data_ptr() == self.leaves[j].data_ptr():
                return j
        return -1


@dataclass
class DeviceTree:
    """Survivor buffer (``WorldTree``, ``world_tree.py:24-351``) resident in HBM."""

    leaves: torch.Tensor      # f64 [n, P+2, 2]
    folds: torch.Tensor       # f32 [n, 2, B]
    scores: torch.Tensor      # f32 [n]
    backtracks: torch.Tensor  # i32 [n, P+2]
    scores_ep: torch.Tensor = field(default=None)

    def __post_init__(self) -> None:
        if self.scores_ep is None:
            self.scores_ep = self.scores.clone()

    @property
    def size(self) -> int:
        return int(self.scores.shape[0])

    @property
    def valid_size(self) -> int:
        return self.size

    @property
    def size_lb(self) -> float:
        return float(np.log2(self.size)) if self.size else 0.0

    @property
    def score_max(self) -> float:
        return float(self.scores.max()) if self.size else 0.0

    @property
    def score_min(self) -> float:
        return float(self.scores.min()) if self.size else 0.0



def circ_child_positions(off1: torch.Tensor, s1_parent: torch.Tensor, s1_expand: torch.Tensor,
                         cr_num: torch.Tensor, batch_size: int) -> tuple[torch.Tensor, int, int]:
    """Final position of every held child's first output (``circular.py:359-394``).

    Per batch of ``batch_size`` parents the children are the held children that do not
    expand, in held order, followed by the ``cr_num[parent]`` crackle children of every
    expanding held child, in held order. Returns (positions, total children, batches)."""
    n = int(off1.shape[0]) - 1
    dev = off1.device
    bs = max(1, min(batch_size, n))
    expand = s1_expand.bool()
    par = s1_parent.long()
    n_s1 = int(par.shape[0])
    nout = torch.where(expand, cr_num.long()[par], torch.ones_like(par))
    ck0 = torch.zeros(n_s1 + 1, dtype=torch.int64, device=dev)
    torch.cumsum((~expand).long(), 0, out=ck0[1:])
    ce0 = torch.zeros(n_s1 + 1, dtype=torch.int64, device=dev)
    torch.cumsum(torch.where(expand, nout, torch.zeros_like(nout)), 0, out=ce0[1:])
    bfirst = torch.arange(0, n, bs, device=dev)
    bstart = off1[bfirst]
    bend = off1[torch.clamp(bfirst + bs, max=n)]
    nk_b = ck0[bend] - ck0[bstart]
    tot_b = nk_b + (ce0[bend] - ce0[bstart])
    base_b = torch.zeros_like(tot_b)
    base_b[1:] = torch.cumsum(tot_b, 0)[:-1]
    b_of = par // bs
    pos = base_b[b_of] + torch.where(expand, nk_b[b_of] + ce0[:-1] - ce0[bstart][b_of],
                                     ck0[:-1] - ck0[bstart][b_of])
    return pos, int(tot_b.sum().item()), int(bfirst.shape[0])

def _raise_branch_error(errv: int) -> None:
    """The branch kernels' error word (smallest failing (parent * P + param) * 4 + kind)."""
    if errv != INT64_MAX:
        if errv % 4 == 1:
            raise ValueError("Invalid input: increase branch_max.")
        raise ValueError("Both dparam_cur and dparam_new must be positive.")


def compact(mask: torch.Tensor, count: int) -> torch.Tensor:
    """Indices of the true entries of ``mask`` in increasing order, ``count`` of them,
    without a host synchronisation: each true entry goes to its prefix-sum slot, every
    false entry to a discarded extra slot."""
    n = int(mask.shape[0])
    pos = torch.cumsum(mask, 0) - 1
    out = torch.empty(count + 1, dtype=torch.int64, device=mask.device)
    out.index_copy_(0, torch.where(mask, pos, torch.full_like(pos, count)),
                    torch.arange(n, dtype=torch.int64, device=mask.device))
    return out[:count]


def overflow_survivors(scores: np.ndarray, child_off: np.ndarray, batch_size: int,
                       threshold: float, max_sugg: int) -> np.ndarray:
    """Children kept by the reference's batch-sequential loop when the buffer overflows.

    Replays ``pruning_iteration_batched`` (``prune.py:214-321``) on the child scores only:
    batches of ``batch_size`` parents, ``max(threshold, current_threshold)`` per batch, and
    ``WorldTree.add_batch`` / ``prune_on_overload`` (``world_tree.py:482-551``). Returns the
    surviving child indices in increasing order.
    """
    n_par = len(child_off) - 1
    bs = max(1, min(batch_size, n_par))
    kept_idx = np.empty(0, dtype=np.int64)
    kept_sc = np.empty(0, dtype=np.float32)
    cur = threshold
    for b0 in range(0, n_par, bs):
        b1 = min(b0 + bs, n_par)
        c0, c1 = int(child_off[b0]), int(child_off[b1])
        if c1 == c0:
            continue
        sc = scores[c0:c1]
        sel = np.where(sc.astype(np.float64) >= max(threshold, cur))[0]
        if len(sel) == 0:
            continue
        bsc = sc[sel]
        bidx = sel.astype(np.int64) + c0
        space = max_sugg - len(kept_idx)
        if len(bsc) <= space:
            kept_idx = np.concatenate([kept_idx, bidx])
            kept_sc = np.concatenate([kept_sc, bsc])
            continue
        total = len(kept_sc) + len(bsc)
        scratch = np.concatenate([kept_sc, bsc])
        k_idx = -max_sugg
        mid_idx = -(total // 2) - 1
        part = np.partition(scratch, [k_idx, mid_idx])
        topk = math.nextafter(float(part[k_idx]), math.inf)
        med = math.nextafter(float(part[mid_idx]), math.inf)
        eff = max(cur, topk, med)
        keep = kept_sc.astype(np.float64) >= eff
        kept_idx, kept_sc = kept_idx[keep], kept_sc[keep]
        pend = np.where(bsc.astype(np.float64) >= eff)[0]
        m = min(len(pend), max_sugg - len(kept_idx))
        if m > 0:
            kept_idx = np.concatenate([kept_idx, bidx[pend[:m]]])
            kept_sc = np.concatenate([kept_sc, bsc[pend[:m]]])
        cur = eff
    return kept_idx


class Pruning:
    """EP search from one reference segment; API of ``prune.py:347-755``."""

    def __init__(self, dyp, threshold_scheme: np.ndarray, max_sugg: int = 2**17,
                 batch_size: int = 1024, poly_basis: str = "taylor", intermediate_k: int = 1,
                 *, use_moving_grid: bool = True, logger=None,  # noqa: ARG002
                 device_loop: bool = True) -> None:
        cfg = dyp.cfg
        self._circular = cfg.prune_poly_order == 5  # prune.py:369-392: circular plugin
        # (is_circular, poly_basis) pairs with a plugin: circular Taylor, polynomial Taylor,
        # polynomial Chebyshev (prune.py:373-381).
        if poly_basis not in ("taylor", "chebyshev") or (self._circular
                                                         and poly_basis == "chebyshev"):
            msg = f"Invalid pruning configuration: {(self._circular, poly_basis, cfg.use_fourier)}"
            raise ValueError(msg)
        self._cheby = poly_basis == "chebyshev"
        self._complex = bool(cfg.use_fourier)  # complex64 folds (Fourier domain)
        if cfg.nparams != cfg.prune_poly_order or cfg.nparams < 2:
            msg = "Taylor pruning needs prune_poly_order == number of FFA parameters (>= 2)"
            raise ValueError(msg)
        self._dyp = dyp
        self._threshold_scheme = np.asarray(threshold_scheme, dtype=np.float64)
        self._max_sugg = int(max_sugg)
        self._batch_size = int(batch_size)
        self._intermediate_k = intermediate_k
        self._use_moving_grid = use_moving_grid
        self._P = cfg.prune_poly_order
        counts = [len(p) for p in dyp.param_arr]
        widths = cfg.score_widths
        pp = PruneParams()
        pp.poly_order = self._P
        pp.nbins = cfg.nbins
        pp.branch_max = cfg.branch_max
        pp.tiling = _TILING[cfg.tiling_strategy]
        pp.eta = cfg.eta
        pp.nparams_ffa = cfg.nparams
        pp.n_accel = counts[-2]
        pp.n_freq = counts[-1]
        pp.nw = len(widths)
        pp.accel_lim[0], pp.accel_lim[1] = (float(x) for x in cfg.param_limits[-2])
        pp.freq_lim[0], pp.freq_lim[1] = (float(x) for x in cfg.param_limits[-1])
        for i, w in enumerate(widths.tolist()):
            pp.widths[i] = w
        pp.fixed_grid = 0 if use_moving_grid else 1
        self._pp = pp
        self._qq = None
        if self._circular:
            qq = CircParams()
            qq.prop_sig = float(cfg.propagator_significance)
            qq.val_sig = float(cfg.validation_significance)
            qq.omega_max_sq = float((2 * np.pi / cfg.p_orb_min) ** 2)
            qq.x_mass_const = float(cfg.x_mass_const)
            qq.batch_size = self._batch_size
            self._qq = qq
        self._widths = widths
        self._ncoords = int(np.prod(counts))
        # Called as level_hook(self, tree) -> tree after every level (no reference
        # counterpart): lets a budgeted search thin the survivors before the next level.
        self.level_hook = None
        # With timing on, HIP events on the launch stream bracket every library call
        # (the GPU side of the reference's per-level timers, prune.py:217-322).
        self.timing = False
        self._events: list[tuple[str, torch.cuda.Event, torch.cuda.Event]] = []
        # execute() runs the levels through the sync-free loop (pyl_ep_levels) when this is
        # set and no Python level hook is installed; execute_iter() is always per level.
        self.device_loop = bool(device_loop)
        # Capacities of the loop's buffers (None: defaults from max_sugg, see _loop_buffers).
        self.loop_caps: dict[str, int] | None = None
        self._loop: LoopBuffers | None = None
        self._loop_ms: list[tuple[str, float]] = []
        self._timer = None
        # Levels the loop handed to the host, by status (ST_*).
        self.loop_handoffs: dict[int, int] = {}
        # (status, survivors before any budget) of every level the loop ran, by level.
        self.loop_levels: dict[int, tuple[int, int]] = {}

    def _timed(self, phase: str, fn, *args):
        if not self.timing:
            return fn(*args)
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        rc = fn(*args)
        b.record()
        self._events.append((phase, a, b))
        return rc

    def kernel_times(self) -> dict[str, list[float]]:
        """Milliseconds of every timed library call, per phase, in call order (the level
        loop's phases are timed inside the loop: branch, children, select, materialize)."""
        torch.cuda.synchronize()
        out: dict[str, list[float]] = {}
        for phase, a, b in self._events:
            out.setdefault(phase, []).append(a.elapsed_time(b))
        for phase, ms in self._loop_ms:
            out.setdefault(phase, []).append(ms)
        return out

    def _loop_timer(self, lib, buf: LoopBuffers, nlev: int) -> None:
        """Attach (timing on) or detach the loop's phase timer."""
        if not self.timing:
            buf.s.timer = None
            return
        if self._timer is None or self._timer[1] < 4 * nlev:
            if self._timer is not None:
                lib.pyl_ep_timer_destroy(self._timer[0])
            h = ctypes.c_void_p()
            _lib.check(lib.pyl_ep_timer_create(4 * nlev, ctypes.byref(h)))
            self._timer = (h, 4 * nlev)
        buf.s.timer = self._timer[0]

    def _read_loop_timer(self, lib) -> None:
        if not self.timing or self._timer is None:
            return
        cap = self._timer[1]
        ph = (ctypes.c_int32 * cap)()
        ms = (ctypes.c_float * cap)()
        n = ctypes.c_int32()
        _lib.check(lib.pyl_ep_timer_read(self._timer[0], ph, ms, cap, ctypes.byref(n)))
        self._loop_ms.extend((EP_PHASES[ph[i]], float(ms[i])) for i in range(n.value))

    def __del__(self) -> None:
        t = getattr(self, "_timer", None)
        if t is not None:
            try:
                _lib.load().pyl_ep_timer_destroy(t[0])
            except Exception:  # noqa: BLE001 - interpreter shutdown
                pass

    @property
    def dyp(self):
        return self._dyp

    @property
    def threshold_scheme(self) -> np.ndarray:
        return self._threshold_scheme

    @property
    def max_sugg(self) -> int:
        return self._max_sugg

    @property
    def batch_size(self) -> int:
        return self._batch_size

    @property
    def prune_level(self) -> int:
        return self._prune_level

    @property
    def prune_complete(self) -> bool:
        return self._prune_complete

    @property
    def scheme(self) -> MiddleOutScheme:
        return self._scheme

    @property
    def world_tree(self) -> DeviceTree:
        return self._world_tree

    @property
    def pstats(self) -> PruneStatsCollection:
        return self._pstats

    @property
    def backtrack_hist(self) -> list[torch.Tensor]:
        return self._backtrack_hist

    def _segment_rows(self, seg_idx: int) -> torch.Tensor:
        fold = self.dyp.fold_device
        return fold[seg_idx]

    def initialize(self, ref_seg: int, log_file: Path | None = None) -> None:
        """Seed the tree from the reference segment (``prune.py:468-524``)."""
        dyp = self.dyp
        self._scheme = MiddleOutScheme(dyp.nsegments, ref_seg, dyp.tseg)
        self._prune_complete = False
        self._prune_level = 0
        P = self._P
        if self._cheby:
            leaves = poly_chebyshev_seed(dyp.param_arr, dyp.dparams_actual, P,
                                         self.scheme.get_coord(0))
        else:
            leaves = poly_taylor_seed(dyp.param_arr, dyp.dparams_actual, P)
        n = leaves.shape[0]
        folds = self._segment_rows(self.scheme.ref_idx).reshape(n, 2, -1).contiguous()
        from pyloki_amd import kernels

        if self._complex:
            scores = kernels.score_boxcar_complex(folds, self._widths, self.dyp.cfg.nbins,
                                                  mode=_lib.PYL_SCORE_MAX)
        else:
            scores = kernels.score_boxcar(folds, self._widths, mode=_lib.PYL_SCORE_MAX)
        bt = torch.zeros((n, P + 2), dtype=torch.int32, device=folds.device)
        self._world_tree = DeviceTree(to_device(leaves), folds.clone(), scores, bt)
        self._backtrack_hist = []
        self._pstats = PruneStatsCollection()
        self._pstats.append(PruneStats(level=0, seg_idx=self.scheme.ref_idx, threshold=0,
                                       score_min=self._world_tree.score_min,
                                       score_max=self._world_tree.score_max, n_branches=n,
                                       n_leaves=n, n_leaves_phy=n, n_leaves_surv=n))
        if log_file is not None:
            with Path(log_file).open("a") as f:
                f.write(self._pstats[0].get_summary())

    def _coords(self, level: int) -> PruneCoords:
        s = self.scheme
        c = PruneCoords()
        mg = self._use_moving_grid
        c.t0_init = s.get_coord(0)[0]
        c.t0_cur, c.scale_cur = s.get_current_coord(level, mg)
        c.t0_next, c.scale_next = s.get_coord(level)
        c.t0_add = s.get_segment_coord(level)[0]
        c.t0_prev, c.scale_prev = s.get_previous_coord(level, mg)
        return c

    def execute_iter(self, log_file: Path | None = None) -> None:
        """One pruning level (``prune.py:615-669``)."""
        if self._prune_complete:
            return
        self._prune_level += 1
        level = self._prune_level
        seg_idx = self.scheme.get_segment_idx(level)
        threshold = float(self.threshold_scheme[level - 1])
        tree = self._world_tree
        new_tree, stats = self._level(tree, self._segment_rows(seg_idx), self._coords(level),
                                      threshold)
        if self.level_hook is not None and new_tree.size:
            new_tree = self.level_hook(self, new_tree)  # e.g. a learned budget (cnn.py)
        st = PruneStats(level=level, seg_idx=seg_idx, threshold=threshold,
                        n_branches=tree.size, n_leaves_surv=new_tree.size, **stats)
        self._end_level(st, new_tree, log_file, new_tree.backtracks.clone())

    def _end_level(self, st: PruneStats, new_tree: DeviceTree, log_file: Path | None,
                   backtracks: torch.Tensor) -> None:
        self._prune_level = st.level
        self._pstats.append(st)
        if log_file is not None:
            with Path(log_file).open("a") as f:
                f.write(st.get_summary())
        self._world_tree = new_tree
        if new_tree.size == 0:
            self._prune_complete = True
            return
        self._backtrack_hist.append(backtracks)

    def _use_device_loop(self) -> bool:
        hook = self.level_hook
        return self.device_loop and (hook is None or hasattr(hook, "device_budget"))

    def _loop_buffers(self, n_in: int, nlev: int) -> LoopBuffers:
        """The loop's buffers, (re)allocated when too small. Defaults: tree buffers of
        max(max_sugg, input) rows, 4 * that many children per level, a history of
        16 * max_sugg rows (at least one level's survivors; a level starts only while a
        full survivor buffer still fits)."""
        caps = self.loop_caps or {}
        cap_par = max(self.max_sugg, n_in)
        buf = self._loop
        if buf is None or buf.cap_par < cap_par or buf.nlev < nlev:
            cap_child = int(caps.get("cap_child", max(2**16, 4 * cap_par)))
            cap_hist = max(self.max_sugg, int(caps.get("cap_hist", 16 * self.max_sugg)))
            buf = LoopBuffers(self._P, self._new_folds, cap_par, cap_child, cap_hist,
                              self.max_sugg, max(nlev, self.dyp.nsegments), self._cheby,
                              self._circular)
            self._loop = buf
        return buf

    def execute_levels(self, last_level: int | None = None,
                       log_file: Path | None = None) -> None:
        """Prune levels prune_level + 1 .. last_level (default: every level left): through
        the sync-free loop when ``device_loop`` is set and no level hook is installed,
        else level by level."""
        nlast = self.dyp.nsegments - 1
        last = nlast if last_level is None else min(int(last_level), nlast)
        if self._use_device_loop():
            self._run_levels(last, log_file)
            return
        while not self._prune_complete and self._prune_level < last:
            self.execute_iter(log_file)

    def _run_levels(self, last: int, log_file: Path | None = None) -> None:
        """Levels prune_level + 1 .. last through the sync-free loop (one host transfer
        per call, for the level records). A level the loop stops at (branch error, child
        buffer, survivor overflow) runs through execute_iter; a full history is replaced."""
        while not self._prune_complete and self._prune_level < last:
            run = self._enqueue_levels(self._prune_level + 1, last, self._world_tree)
            self._finish_levels(run, log_file)

    def _enqueue_levels(self, l0: int, last: int, tree: DeviceTree) -> dict:
        """Enqueue levels l0 .. last on the current stream, reading ``tree``; no host
        synchronisation (a graph may capture it). Returns what ``_finish_levels`` needs."""
        lib = _lib.load()
        fold = self.dyp.fold_device
        stride = int(fold[0].numel()) * (2 if fold.is_complex() else 1)  # floats per segment
        flags = (EP_COMPLEX if self._complex else 0) | (EP_CHEBYSHEV if self._cheby else 0)
        levels = range(l0, last + 1)
        nlev = len(levels)
        buf = self._loop_buffers(tree.size, nlev)
        coords = (PruneCoords * nlev)(*[self._coords(lv) for lv in levels])
        thr = np.ascontiguousarray(self.threshold_scheme[l0 - 1:last], dtype=np.float64)
        segs = np.array([self.scheme.get_segment_idx(lv) for lv in levels], dtype=np.int64)
        budget = None
        if self.level_hook is not None:  # a survivor budget the loop applies itself
            if buf.holds(tree) == 0:
                tree = DeviceTree(tree.leaves.clone(), tree.folds.clone(), tree.scores.clone(),
                                  tree.backtracks)
            budget = self.level_hook.device_budget(self, buf.cap_par, l0)
        buf.s.budget = ctypes.addressof(budget) if budget is not None else None
        first_out = 1 if buf.holds(tree) == 0 else 0
        self._loop_timer(lib, buf, nlev)
        args = (coords, thr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                segs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), nlev, ptr(fold), stride,
                ptr(tree.leaves), ptr(tree.folds), tree.size, buf.hist_next, first_out,
                ctypes.byref(buf.s))
        if self._circular:
            _lib.check(self._timed(
                "levels", lib.pyl_circ_levels, ctypes.byref(self._pp), ctypes.byref(self._qq),
                flags, *args, ctypes.byref(buf.c), stream_ptr()))
        else:
            _lib.check(self._timed("levels", lib.pyl_ep_levels, ctypes.byref(self._pp), flags,
                                   *args, stream_ptr()))
        return {"levels": levels, "buf": buf, "thr": thr, "segs": segs, "budget": budget,
                "first_out": first_out}

    def _finish_levels(self, run: dict, log_file: Path | None = None) -> int:
        """Read the records of an enqueued run (one transfer) and advance the search over
        its finished levels; a level the loop stopped at is run through execute_iter (or
        the history is replaced). Returns the stop status."""
        lib = _lib.load()
        levels, buf, thr, segs = run["levels"], run["buf"], run["thr"], run["segs"]
        budget, first_out = run["budget"], run["first_out"]
        nlev = len(levels)
        self._read_loop_timer(lib)
        # one transfer: the records, then the min/max pairs' bits as int64 words
        info = torch.cat([buf.rec[:nlev + 1].reshape(-1),
                          buf.minmax[:nlev].view(torch.int64).reshape(-1)]).cpu().numpy()
        rec = info[:(nlev + 1) * REC_WORDS].reshape(nlev + 1, REC_WORDS)
        mm = info[(nlev + 1) * REC_WORDS:].view(np.float32).reshape(nlev, 2)
        stop = ST_OK
        for i, level in enumerate(levels):
            r = rec[i]
            stop = int(r[REC_STATUS])
            self.loop_levels[level] = (stop, int(r[REC_NSURV]))
            if stop != ST_OK:
                break
            o = 1 if budget is not None else (first_out + i) & 1
            m, h, nv = int(r[REC_NSURV]), int(r[REC_HIST]), int(r[REC_NVALID])
            if budget is not None:
                m = int(r[REC_NKEPT])
            sc = buf.scores[o][:m]
            new_tree = DeviceTree(buf.leaves[o][:m], buf.folds[o][:m], sc, buf.hist[h:h + m],
                                  scores_ep=sc)
            st = PruneStats(level=level, seg_idx=int(segs[i]), threshold=float(thr[i]),
                            score_min=float(mm[i, 0]) if nv else 0.0,
                            score_max=float(mm[i, 1]) if nv else 0.0,
                            n_branches=int(r[REC_NPAR]), n_leaves=int(r[REC_NCHILD]),
                            n_leaves_phy=nv, n_leaves_surv=m)
            buf.hist_next = h + m
            self._end_level(st, new_tree, log_file, new_tree.backtracks)
            if self._prune_complete:
                break
        if stop == ST_OK or self._prune_complete:
            return stop
        self.loop_handoffs[stop] = self.loop_handoffs.get(stop, 0) + 1
        if stop == ST_HIST_FULL:
            buf.new_hist(2 * int(buf.s.cap_hist))
            return stop
        self.execute_iter(log_file)  # raises on a branch error; replays an overflow
        if stop == ST_CHILD_CAP:
            buf.set_children(2 * max(buf.cap_child, self._pstats[-1].n_leaves))
        return stop

    def _fn(self, lib, name: str):
        """The real or complex-fold variant of an EP entry point."""
        return getattr(lib, f"{name}_complex" if self._complex else name)

    def _new_folds(self, n: int, dev) -> torch.Tensor:
        B = self.dyp.cfg.nbins
        if self._complex:
            return torch.empty((n, 2, B // 2 + 1), dtype=torch.complex64, device=dev)
        return torch.empty((n, 2, B), dtype=torch.float32, device=dev)

    def _select(self, c_score: torch.Tensor, minmax: torch.Tensor, child_off_fn,
                threshold: float) -> tuple[torch.Tensor, float, float]:
        """Children kept and the level's score range, with one host transfer: all passing
        children when they fit the buffer, else the reference's batch-sequential overflow
        rule replayed on the scores (``child_off_fn()`` gives the per-parent child offsets
        that mark the batch boundaries)."""
        passing = c_score.double() >= threshold
        info = torch.cat([passing.sum().reshape(1).double(), minmax.double()]).cpu()
        npass = int(info[0])
        smin, smax = float(info[1]), float(info[2])
        if npass <= self.max_sugg:
            return compact(passing, npass), smin, smax
        surv_np = overflow_survivors(c_score.cpu().numpy(), child_off_fn(), self.batch_size,
                                     threshold, self.max_sugg)
        return torch.from_numpy(surv_np).to(c_score.device), smin, smax

    def _level(self, tree: DeviceTree, seg: torch.Tensor, coords: PruneCoords,
               threshold: float) -> tuple[DeviceTree, dict]:
        if self._circular:
            return self._level_circ(tree, seg, coords, threshold)
        lib = _lib.load()
        dev = tree.folds.device
        P = self._P
        n = tree.size
        pp, cc, st = ctypes.byref(self._pp), ctypes.byref(coords), stream_ptr()
        branch = torch.empty((n, P, 2), dtype=torch.float64, device=dev)
        counts = torch.empty((n, P), dtype=torch.int32, device=dev)
        nchild = torch.empty(n, dtype=torch.int64, device=dev)
        err = torch.full((1,), INT64_MAX, dtype=torch.int64, device=dev)
        if self._cheby:
            # the parents moved to coord_cur feed the children (chebyshev.py:112-119)
            src = torch.empty_like(tree.leaves)
            _lib.check(self._timed("branch", lib.pyl_cheb_branch, pp, cc, ptr(tree.leaves), n,
                                   ptr(src), ptr(branch), ptr(counts), ptr(nchild), ptr(err),
                                   st))
            pre = "pyl_cheb"
        else:
            src = tree.leaves
            _lib.check(self._timed("branch", lib.pyl_prune_branch, pp, cc, ptr(tree.leaves), n,
                                   ptr(branch), ptr(counts), ptr(nchild), ptr(err), st))
            pre = "pyl_prune"
        off = torch.zeros(n + 1, dtype=torch.int64, device=dev)
        torch.cumsum(nchild, 0, out=off[1:])
        errv, total = (int(x) for x in torch.stack([err[0], off[-1]]).cpu())
        _raise_branch_error(errv)
        c_parent = torch.empty(total, dtype=torch.int32, device=dev)
        c_row = torch.empty(total, dtype=torch.int32, device=dev)
        c_shift = torch.empty(total, dtype=torch.int32, device=dev)
        c_phase = torch.empty(total, dtype=torch.float64, device=dev)
        c_score = torch.empty(total, dtype=torch.float32, device=dev)
        minmax = torch.tensor([np.inf, -np.inf], dtype=torch.float32, device=dev)
        segc = seg.contiguous()
        _lib.check(self._timed(
            "children", self._fn(lib, f"{pre}_children"), pp, cc, ptr(src), ptr(tree.folds), n,
            ptr(branch), ptr(counts), ptr(off), total, ptr(segc), ptr(c_parent), ptr(c_row),
            ptr(c_shift), ptr(c_phase), ptr(c_score), ptr(minmax), st))
        surv, smin, smax = self._select(c_score, minmax, lambda: off.cpu().numpy(), threshold)
        m = int(surv.shape[0])
        o_leaves = torch.empty((m, P + 2, 2), dtype=torch.float64, device=dev)
        o_folds = self._new_folds(m, dev)
        o_scores = torch.empty(m, dtype=torch.float32, device=dev)
        o_bt = torch.empty((m, P + 2), dtype=torch.int32, device=dev)
        _lib.check(self._timed(
            "materialize", self._fn(lib, f"{pre}_materialize"), pp, cc, ptr(src),
            ptr(tree.folds), ptr(branch), ptr(counts), ptr(off), ptr(segc), ptr(c_parent),
            ptr(c_row), ptr(c_shift), ptr(c_phase), ptr(c_score), ptr(surv.contiguous()), m,
            ptr(o_leaves), ptr(o_folds), ptr(o_scores), ptr(o_bt), st))
        stats = {"n_leaves": total, "n_leaves_phy": total,
                 "score_min": smin if total else 0.0, "score_max": smax if total else 0.0}
        return DeviceTree(o_leaves, o_folds, o_scores, o_bt), stats

    def _level_circ(self, tree: DeviceTree, seg: torch.Tensor, coords: PruneCoords,
                    threshold: float) -> tuple[DeviceTree, dict]:
        """One circular-basis level (``circular.py`` via ``prune.py:165-332``). Four host
        transfers size the buffers: held children, final children, validated children,
        survivors."""
        lib = _lib.load()
        dev = tree.folds.device
        P = self._P
        n = tree.size
        bs = max(1, min(self.batch_size, n))
        pp, qq, cc = ctypes.byref(self._pp), ctypes.byref(self._qq), ctypes.byref(coords)
        st = stream_ptr()
        i64, i32 = torch.int64, torch.int32
        buf = {
            "branch": torch.empty((n, P, 2), dtype=torch.float64, device=dev),
            "counts": torch.empty((n, P), dtype=i32, device=dev),
            "cr_first": torch.empty(n, dtype=torch.float64, device=dev),
            "cr_num": torch.empty(n, dtype=i32, device=dev),
        }
        L = CircLevel()
        for name, t in buf.items():
            setattr(L, name, ptr(t))
        L.n_par = n
        nchild1 = torch.empty(n, dtype=i64, device=dev)
        err = torch.full((1,), INT64_MAX, dtype=i64, device=dev)
        _lib.check(self._timed("branch", lib.pyl_circ_branch, pp, qq, cc, ptr(tree.leaves),
                               ctypes.byref(L), ptr(nchild1), ptr(err), st))
        off1 = torch.zeros(n + 1, dtype=i64, device=dev)
        torch.cumsum(nchild1, 0, out=off1[1:])
        errv, n_s1 = (int(x) for x in torch.stack([err[0], off1[-1]]).cpu())
        _raise_branch_error(errv)
        buf["off1"] = off1
        buf["s1_parent"] = torch.empty(n_s1, dtype=i32, device=dev)
        buf["s1_expand"] = torch.empty(n_s1, dtype=torch.uint8, device=dev)
        for name in ("off1", "s1_parent", "s1_expand"):
            setattr(L, name, ptr(buf[name]))
        L.n_s1 = n_s1
        _lib.check(self._timed("branch", lib.pyl_circ_stage1, pp, qq, cc, ptr(tree.leaves),
                               ctypes.byref(L), st))
        pos, total, nb = circ_child_positions(off1, buf["s1_parent"], buf["s1_expand"],
                                              buf["cr_num"], bs)
        buf["s1_pos"] = pos.contiguous()
        buf["c_s1"] = torch.empty(total, dtype=i64, device=dev)
        buf["c_sub"] = torch.empty(total, dtype=i32, device=dev)
        buf["c_keep"] = torch.empty(total, dtype=torch.uint8, device=dev)
        buf["c_region"] = torch.empty(total, dtype=torch.int8, device=dev)
        for name in ("s1_pos", "c_s1", "c_sub", "c_keep", "c_region"):
            setattr(L, name, ptr(buf[name]))
        L.n_child = total
        _lib.check(self._timed("branch", lib.pyl_circ_expand, pp, qq, cc, ptr(tree.leaves),
                               ctypes.byref(L), st))
        keep = buf["c_keep"].bool()
        nv = int(keep.sum().item())
        valid = compact(keep, nv)
        c_parent = torch.empty(nv, dtype=i32, device=dev)
        c_row = torch.empty(nv, dtype=i32, device=dev)
        c_shift = torch.empty(nv, dtype=i32, device=dev)
        c_phase = torch.empty(nv, dtype=torch.float64, device=dev)
        c_score = torch.empty(nv, dtype=torch.float32, device=dev)
        minmax = torch.tensor([np.inf, -np.inf], dtype=torch.float32, device=dev)
        segc = seg.contiguous()
        _lib.check(self._timed(
            "children", self._fn(lib, "pyl_circ_children"), pp, qq, cc, ptr(tree.leaves),
            ptr(tree.folds), ctypes.byref(L), ptr(valid), nv, ptr(segc), ptr(c_parent),
            ptr(c_row), ptr(c_shift), ptr(c_phase), ptr(c_score), ptr(minmax), st))

        def valid_off() -> np.ndarray:
            cnt = torch.bincount(c_parent.long(), minlength=n)
            return np.concatenate([[0], np.cumsum(cnt.cpu().numpy())])

        surv, smin, smax = self._select(c_score, minmax, valid_off, threshold)
        surv = surv.contiguous()
        m = int(surv.shape[0])
        # Batches whose validated children include a Taylor cell (circular.py:506-511).
        tay = (buf["c_region"][valid] == REGION_TAYLOR).to(i32)
        ht = torch.zeros(nb, dtype=i32, device=dev)
        ht.scatter_reduce_(0, c_parent.long() // bs, tay, reduce="amax")
        has_taylor = ht.to(torch.uint8)
        o_leaves = torch.empty((m, P + 2, 2), dtype=torch.float64, device=dev)
        o_folds = self._new_folds(m, dev)
        o_scores = torch.empty(m, dtype=torch.float32, device=dev)
        o_bt = torch.empty((m, P + 2), dtype=i32, device=dev)
        _lib.check(self._timed(
            "materialize", self._fn(lib, "pyl_circ_materialize"), pp, qq, cc,
            ptr(tree.leaves), ptr(tree.folds), ctypes.byref(L), ptr(valid), ptr(segc),
            ptr(c_parent), ptr(c_row), ptr(c_shift), ptr(c_phase), ptr(c_score),
            ptr(has_taylor), ptr(surv), m, ptr(o_leaves), ptr(o_folds), ptr(o_scores),
            ptr(o_bt), st))
        stats = {"n_leaves": total, "n_leaves_phy": nv,
                 "score_min": smin if nv else 0.0, "score_max": smax if nv else 0.0}
        return DeviceTree(o_leaves, o_folds, o_scores, o_bt), stats

    def _do_ascend(self) -> None:
        """Re-integrate the survivors over all segments so far (``prune.py:671-693``)."""
        if self._prune_complete or self._world_tree.size == 0:
            return
        tree = self._world_tree
        seg_idx, seg_coords = self.scheme.get_segment_coords_so_far(self._prune_level)
        # coord_end (prune.py:680-683): where the survivors' parameters are referenced.
        t0_cur, scale_cur = self.scheme.get_previous_coord(self._prune_level + 1,
                                                           self._use_moving_grid)
        n = tree.size
        folds = self._new_folds(n, tree.folds.device)
        scores = torch.empty(n, dtype=torch.float32, device=tree.folds.device)
        idx = np.ascontiguousarray(seg_idx, dtype=np.int64)
        t0 = np.ascontiguousarray(seg_coords[:, 0], dtype=np.float64)
        args = (ptr(tree.leaves), n, ptr(self.dyp.fold_device.contiguous()), self._ncoords,
                idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                t0.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(idx), t0_cur,
                ptr(folds), ptr(scores), stream_ptr())
        lib = _lib.load()
        if self._circular:
            _lib.check(self._timed("ascend", self._fn(lib, "pyl_circ_ascend"),
                                   ctypes.byref(self._pp), ctypes.byref(self._qq), *args))
        elif self._cheby:
            _lib.check(self._timed("ascend", self._fn(lib, "pyl_cheb_ascend"),
                                   ctypes.byref(self._pp), *args[:8], scale_cur, *args[8:]))
        else:
            _lib.check(self._timed("ascend", self._fn(lib, "pyl_prune_ascend"),
                                   ctypes.byref(self._pp), *args))
        tree.scores_ep = tree.scores.clone()
        tree.folds = folds
        tree.scores = scores

    def execute(self, ref_seg: int, ascend_levels: list[int] | None = None,
                outdir: str | Path = "./", log_file: Path | None = None,  # noqa: ARG002
                result_file: Path | None = None, shared_progress=None,  # noqa: ANN001, ARG002
                task_id: int | None = None) -> dict:
        """Seed, prune every level, ascend, report (``prune.py:526-613``).

        The run is named ``f"{ref_seg:03d}_{task_id:02d}"``; with ``result_file`` it is
        appended to that results file (``PruneResultWriter``) and with ``log_file`` the
        level summaries go there. (The reference writes temporary files in ``outdir``
        when neither is given; here nothing is written then.) The run is also returned:
        run name, scheme, reported parameters, scores, EP scores, level statistics."""
        asc = frozenset(ascend_levels) if ascend_levels is not None else None
        if log_file is not None:
            with Path(log_file).open("a") as f:
                f.write(f"Pruning log for ref segment: {ref_seg}\n")
        self.initialize(ref_seg, log_file)
        nlast = self.dyp.nsegments - 1
        if self._use_device_loop():
            stops = sorted(lv for lv in (asc or ()) if 1 <= lv < nlast)
            for stop in [*stops, nlast]:
                self._run_levels(stop, log_file)
                if asc is not None and self._prune_level == stop and stop in asc:
                    self._do_ascend()
        else:
            for _i in range(nlast):
                self.execute_iter(log_file)
                if asc is not None and self._prune_level in asc:
                    self._do_ascend()
        if asc is None or self._prune_level not in asc:
            self._do_ascend()
        res = self.report(ref_seg, task_id=0 if task_id is None else task_id)
        if result_file is not None:
            with PruneResultWriter(result_file, mode="a") as writer:
                writer.write_run_results(res["run_name"], res["scheme"], res["param_sets"],
                                         res["scores"], res["scores_ep"], self._pstats)
        if log_file is not None:
            with Path(log_file).open("a") as f:
                f.write(f"Pruning run complete for ref segment: {ref_seg}\n")
                f.write(f"Time: {self._pstats.get_timer_summary()}\n\n")
        return res

    def report(self, ref_seg: int, task_id: int = 0) -> dict:
        """The result of a finished search: report transform of the survivors, their
        scores before and after the ascend, and the level statistics (``prune.py:592-613``)."""
        leaves = self._world_tree.leaves.cpu().numpy()
        if not len(leaves):
            report = leaves
        elif self._cheby:
            report = poly_chebyshev_report_batch(
                leaves, self.scheme.get_report_coord(self._use_moving_grid))
        else:
            report = poly_taylor_report_batch(leaves)
        return {
            "run_name": f"{ref_seg:03d}_{task_id:02d}",
            "scheme": self.scheme.data.copy(),
            "param_sets": report,
            "scores": self._world_tree.scores.cpu().numpy(),
            "scores_ep": self._world_tree.scores_ep.cpu().numpy(),
            "level_stats": self._pstats,
        }


def poly_taylor_seed(param_arr: list[np.ndarray], dparams: np.ndarray,
                     poly_order: int) -> np.ndarray:
    """Seed leaves from the last FFA grid (``taylor.py:14-65``, ``common.py:14-33``):
    cartesian product (first parameter slowest), ``v = 0`` with width ``df C / f0``,
    ``d0 = 0``, then (f0, flag 0)."""
    mesh = np.meshgrid(*param_arr, indexing="ij")
    cart = np.stack([m.reshape(-1) for m in mesh], axis=1)
    n = cart.shape[0]
    P = poly_order
    leaves = np.zeros((n, P + 2, 2), dtype=np.float64)
    for k in range(P - 1):
        leaves[:, k, 0] = cart[:, k]
        leaves[:, k, 1] = dparams[k]
    f0 = cart[:, -1]
    leaves[:, -3, 1] = np.full(n, dparams[-1]) * (C_VAL / f0)
    leaves[:, -1, 0] = f0
    return leaves


def poly_chebyshev_seed(param_arr: list[np.ndarray], dparams: np.ndarray, poly_order: int,
                        coord_init: tuple[float, float]) -> np.ndarray:
    """``chebyshev.py:14-69``: the Taylor seed cells as Chebyshev coefficients on the
    reference segment's domain (values and widths)."""
    leaves = poly_taylor_seed(param_arr, dparams, poly_order)
    leaves[:, :-1] = transforms.taylor_to_cheby_full(leaves[:, :-1], coord_init[1])
    return leaves


def poly_taylor_report_batch(leaves: np.ndarray) -> np.ndarray:
    """Gauge transform to physical units with error propagation (``taylor.py:342-360``)."""
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


def poly_chebyshev_report_batch(leaves: np.ndarray,
                                coord_report: tuple[float, float]) -> np.ndarray:
    """Chebyshev coefficients to physical parameters (``chebyshev.py:371-398``): Taylor
    coefficients on the report domain, d0 into its row, then the Taylor gauge transform."""
    out = leaves.copy()
    d = transforms.cheby_to_taylor_full(leaves[:, :-1, :], coord_report[1])
    vals = d[:, :-2, 0]
    sigs = d[:, :-2, 1]
    out[:, -2, 0] = d[:, -1, 0]
    out[:, -2, 1] = d[:, -1, 1]
    v = d[:, -2, 0]
    dv = d[:, -2, 1]
    f0 = leaves[:, -1, 0]
    s = 1 - v / C_VAL
    out[:, :-3, 0] = vals / s[:, None]
    out[:, :-3, 1] = np.sqrt((sigs / s[:, None]) ** 2
                             + ((vals / (C_VAL * s[:, None] ** 2)) ** 2) * (dv[:, None] ** 2))
    out[:, -3, 0] = f0 * s
    out[:, -3, 1] = f0 * dv / C_VAL
    return out


def determine_ref_segs_pareto(nsegments: int, n_runs: int) -> list[int]:
    """Pareto-optimal reference segments (``np_utils.py:412-443``)."""
    if not (1 <= n_runs <= nsegments):
        msg = f"n_runs must be between 1 and {nsegments}"
        raise ValueError(msg)
    if n_runs == 1:
        return [nsegments // 2]

    def rdiv(num: int, den: int) -> int:
        return (num + den // 2) // den

    margin = rdiv(nsegments - 1, 2 * n_runs)
    lo, hi = margin, nsegments - 1 - margin
    denom = n_runs - 1
    return [rdiv((denom - i) * lo + i * hi, denom) for i in range(n_runs)]


def prune_dyp_tree(dyp, threshold_scheme: np.ndarray, n_runs: int | None = None,
                   ref_segs: list[int] | None = None, ascend_levels: list[int] | None = None,
                   max_sugg: int = 2**18, batch_size: int = 1024, outdir: str | Path = "./",
                   file_prefix: str = "test", poly_basis: str = "taylor", *,
                   use_moving_grid: bool = True, n_workers: int = 4) -> str:
    """Run the EP search from several reference segments and write every run to one results
    file; returns its path (``prune.py:824-932``).

    The reference fans the runs out over ``n_workers`` processes. Here every run uses the
    GPU in turn (one run's levels already fill the device); ``n_workers`` is checked as the
    reference checks it and sets the task ids of the run names (0 for every run with one
    worker, else the run's index), so the file's run names are the reference's.
    """
    if not isinstance(n_workers, (int, np.integer)):
        msg = f"n_workers must be an integer, got {type(n_workers)}"
        raise TypeError(msg)
    if n_workers < 1:
        msg = f"n_workers must be greater than 0, got {n_workers}"
        raise ValueError(msg)
    filebase = f"{file_prefix}_pruning_nstages_{dyp.nsegments}"
    log_file = Path(outdir) / f"{filebase}_log.txt"
    result_file = Path(outdir) / f"{filebase}_results.npz"
    outdir = mkdir_p(outdir, must_not_exist=[log_file.name, result_file.name])
    # n_runs takes precedence over ref_segs
    if n_runs is not None:
        if not 1 <= n_runs <= dyp.nsegments:
            msg = f"n_runs must be between 1 and {dyp.nsegments}, got {n_runs}"
            raise ValueError(msg)
        ref_segs = determine_ref_segs_pareto(dyp.nsegments, n_runs)
    elif ref_segs is None:
        msg = "Either n_runs or ref_segs must be provided"
        raise ValueError(msg)
    n_workers_act = min(n_workers, len(ref_segs))
    log_file.write_text("Pruning log\n")
    with PruneResultWriter(result_file) as writer:
        writer.write_metadata(dyp.cfg.param_names, dyp.nsegments, max_sugg, threshold_scheme)
    pr = Pruning(dyp, threshold_scheme, max_sugg=max_sugg, batch_size=batch_size,
                 poly_basis=poly_basis, use_moving_grid=use_moving_grid)
    for task, ref in enumerate(ref_segs):
        pr.execute(int(ref), ascend_levels=ascend_levels, outdir=outdir, log_file=log_file,
                   result_file=result_file, task_id=0 if n_workers_act == 1 else task)
    return result_file.as_posix()
