"""Learned branch ranking for a budgeted EP search.

This part has no counterpart in the reference. A small 1-D CNN (three circular Conv1d
layers, 32 channels, kernel 5, on a survivor's 32/64/128-bin profile, plus 8 parameter
features) predicts whether an EP survivor is an ancestor of a final candidate. The
budgeted search keeps, at every level, the ``budget`` survivors it ranks highest, so the
levels that follow branch fewer parents; the exhaustive search is the judge (recall =
fraction of its final candidates that the budgeted search also returns).

The network runs and trains only through the HIP library (``pyl_cnn_forward``,
``pyl_cnn_backward``, ``pyl_cnn_adam``); torch provides device memory.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np
import torch

from pyloki_amd import _lib
from pyloki_amd.device import ptr, stream_ptr

NFEAT = 8
_BUDGET_SCORE, _BUDGET_CNN, _BUDGET_HYBRID = 1, 2, 3
_BUDGET_SEL_WORDS, _BUDGET_HIST_BINS, _BUDGET_STATS = 272, 1024, 1032
_EP_SCAN_BLOCKS = 256
_REC_WORDS, _REC_NKEPT = 12, 10
_SHAPES = [("W1", (32, 1, 5)), ("b1", (32,)), ("W2", (32, 32, 5)), ("b2", (32,)),
           ("W3", (32, 32, 5)), ("b3", (32,)), ("Wf1", (32, 40)), ("bf1", (32,)),
           ("Wf2", (32,)), ("bf2", (1,))]


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases, in the library's flat order (FP32)."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in _SHAPES:
        if name.startswith("b"):
            parts.append(np.zeros(shape))
            continue
        fan_in = int(np.prod(shape[1:])) if len(shape) > 1 else 32
        parts.append(rng.normal(0.0, np.sqrt(2.0 / fan_in), shape))
    return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)


def leaf_features(leaves: torch.Tensor, scores: torch.Tensor, level: int,
                  nlevels: int) -> torch.Tensor:
    """[n][8] features of one level's survivors (``pyl_cnn_features``): S/N / 10,
    level / nlevels, the S/N's z-score and rank fraction within the level (the rank from a
    1024-bin histogram of the level's S/N), then log10(1 + cell width) / 10 of the four
    lowest-order branched parameters."""
    lib = _lib.load()
    n = int(leaves.shape[0])
    P = int(leaves.shape[1]) - 2
    dev = leaves.device
    f = torch.zeros((n, NFEAT), dtype=torch.float32, device=dev)
    if n == 0:
        return f
    stats = torch.empty(_BUDGET_STATS, dtype=torch.float64, device=dev)
    shist = torch.empty(2 * _BUDGET_HIST_BINS, dtype=torch.int32, device=dev)
    _lib.check(lib.pyl_cnn_features(ptr(leaves.contiguous()), ptr(scores.contiguous().float()),
                                    n, P, int(level), int(nlevels), ptr(stats), ptr(shist),
                                    ptr(f), stream_ptr()))
    return f


class BranchRanker:
    """Parameters and Adam state of the ranking network on the current GPU."""

    def __init__(self, nbins: int, params: np.ndarray | None = None, seed: int = 0) -> None:
        lib = _lib.load()
        self.nparams = int(lib.pyl_cnn_num_params())
        if lib.pyl_cnn_cache_floats(nbins) < 0:
            msg = f"the ranking network takes 32, 64 or 128 bins (got {nbins})"
            raise ValueError(msg)
        self.nbins = int(nbins)
        p = init_params(seed) if params is None else np.asarray(params, dtype=np.float32)
        if p.shape != (self.nparams,):
            msg = f"expected {self.nparams} parameters, got {p.shape}"
            raise ValueError(msg)
        dev = torch.device("cuda")
        self.params = torch.from_numpy(p.copy()).to(dev)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self._work = torch.empty(self.nparams * 2048, dtype=torch.float32, device=dev)
        self.grad = torch.zeros_like(self.params)

    def logits(self, folds: torch.Tensor, feats: torch.Tensor,
               cache: torch.Tensor | None = None, precision: str = "fp32") -> torch.Tensor:
        """The network's logits; ``precision="bf16"`` runs the convolutions on MFMA
        (inference only, no training cache)."""
        lib = _lib.load()
        n = int(folds.shape[0])
        cplx = bool(folds.is_complex())
        out = torch.empty(n, dtype=torch.float32, device=folds.device)
        f = folds.contiguous()
        ft = feats.contiguous().float()
        if precision == "bf16":
            if cache is not None:
                msg = "the BF16 forward pass keeps no training cache"
                raise ValueError(msg)
            _lib.check(lib.pyl_cnn_forward_bf16(ptr(self.params), ptr(f), n, self.nbins,
                                                int(cplx), ptr(ft), ptr(out), stream_ptr()))
        elif precision == "fp32":
            _lib.check(lib.pyl_cnn_forward(ptr(self.params), ptr(f), n, self.nbins, int(cplx),
                                           ptr(ft), ptr(out), ptr(cache), stream_ptr()))
        else:
            msg = f"unknown precision {precision!r}"
            raise ValueError(msg)
        return out

    def train_step(self, folds: torch.Tensor, feats: torch.Tensor, labels: torch.Tensor,
                   lr: float = 1e-3, w_pos: float = 1.0, w_neg: float = 1.0,
                   beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> float:
        """One Adam step on the mean weighted BCE of the batch; returns that loss."""
        lib = _lib.load()
        n = int(folds.shape[0])
        stride = int(lib.pyl_cnn_cache_floats(self.nbins))
        cache = torch.empty(n * stride, dtype=torch.float32, device=folds.device)
        lg = self.logits(folds, feats, cache)
        y = labels.contiguous().float()
        _lib.check(lib.pyl_cnn_backward(ptr(self.params), ptr(cache), ptr(lg), ptr(y), n,
                                        self.nbins, w_pos / n, w_neg / n, ptr(self._work),
                                        int(self._work.numel()), ptr(self.grad),
                                        stream_ptr()))
        self.step += 1
        _lib.check(lib.pyl_cnn_adam(ptr(self.params), ptr(self.grad), ptr(self.m),
                                    ptr(self.v), self.nparams, lr, beta1, beta2, eps,
                                    self.step, stream_ptr()))
        loss = (w_pos * y * torch.nn.functional.softplus(-lg)
                + w_neg * (1 - y) * torch.nn.functional.softplus(lg)).mean()
        return float(loss.item())

    def state(self) -> dict:
        return {"params": self.params.cpu().numpy(), "nbins": self.nbins}

    @classmethod
    def from_state(cls, state: dict) -> BranchRanker:
        return cls(int(state["nbins"]), params=np.asarray(state["params"]))


# ----------------------------------------------------------------------- data from EP runs
@dataclass
class LevelRecord:
    level: int
    folds: torch.Tensor
    feats: torch.Tensor
    backtracks: torch.Tensor


def record_levels(pr, ref_seg: int) -> list[LevelRecord]:
    """Run an exhaustive search level by level, keeping every level's survivors."""
    nlev = pr.dyp.nsegments - 1
    pr.initialize(ref_seg)
    recs = []
    for _ in range(nlev):
        pr.execute_iter()
        t = pr.world_tree
        if t.size == 0:
            break
        recs.append(LevelRecord(pr.prune_level, t.folds.clone(),
                                leaf_features(t.leaves, t.scores, pr.prune_level, nlev),
                                t.backtracks.clone()))
    return recs


def ancestor_labels(recs: list[LevelRecord]) -> list[torch.Tensor]:
    """1 for the survivors of each level whose subtree reaches the last level's tree."""
    labels = [None] * len(recs)
    if not recs:
        return labels
    dev = recs[-1].folds.device
    cur = torch.ones(recs[-1].folds.shape[0], dtype=torch.float32, device=dev)
    labels[-1] = cur
    for k in range(len(recs) - 1, 0, -1):
        prev = torch.zeros(recs[k - 1].folds.shape[0], dtype=torch.float32, device=dev)
        parents = recs[k].backtracks[:, 0].long()
        prev[parents[cur > 0]] = 1.0
        labels[k - 1] = prev
        cur = prev
    return labels


def descendant_counts(recs: list[LevelRecord]) -> list[torch.Tensor]:
    """For the survivors of each level: how many of the last level's survivors (the final
    candidates) descend from it."""
    counts = [None] * len(recs)
    if not recs:
        return counts
    dev = recs[-1].folds.device
    cur = torch.ones(recs[-1].folds.shape[0], dtype=torch.float64, device=dev)
    counts[-1] = cur
    for k in range(len(recs) - 1, 0, -1):
        prev = torch.zeros(recs[k - 1].folds.shape[0], dtype=torch.float64, device=dev)
        prev.index_add_(0, recs[k].backtracks[:, 0].long(), cur)
        counts[k - 1] = prev
        cur = prev
    return counts


def level_needs(recs: list[LevelRecord], weights: list[torch.Tensor], key_fn,
                loss: float = 0.0) -> dict:
    """Per search level of an exhaustive run: the smallest budget whose cut, ranking the
    survivors by ``key_fn(record)`` (ties keep the lowest rows, as the cut does), loses at
    most the fraction ``loss`` of the final candidates (``weights``: each survivor's final
    descendants, ``descendant_counts``; with loss 0 every ancestor is kept)."""
    out = {}
    for r, wt in zip(recs, weights):
        total = float(wt.sum().item())
        if total <= 0:
            out[r.level] = 0
            continue
        key = key_fn(r).float()
        order = torch.argsort(-key, stable=True)
        w_sorted = wt[order]
        # lost(b) = weight ranked at b or below; the smallest b with lost(b) <= loss*total
        tail = torch.flip(torch.cumsum(torch.flip(w_sorted, [0]), 0), [0])  # tail[b]
        ok = torch.nonzero(tail <= loss * total * (1 + 1e-12)).reshape(-1)
        b = int(ok[0].item()) if len(ok) else len(order)
        out[r.level] = b
    return out


def budget_schedule(needs: list[dict], nlevels: int, margin: float = 1.5, floor: int = 512,
                    window: int = 2) -> np.ndarray:
    """A budget per search level (0..nlevels) from the needs of calibration runs: the
    largest need of any run within ``window`` levels, times ``margin``, at least
    ``floor``."""
    need = np.zeros(nlevels + 1, dtype=np.float64)
    for nd in needs:
        for level, v in nd.items():
            need[level] = max(need[level], v)
    sched = np.empty_like(need)
    for level in range(nlevels + 1):
        lo, hi = max(0, level - window), min(nlevels, level + window)
        sched[level] = need[lo:hi + 1].max()
    return np.maximum(np.ceil(margin * sched), floor).astype(np.int64)


def snr_key(r: LevelRecord) -> torch.Tensor:
    return r.feats[:, 0]


def network_keys(ranker: BranchRanker, hybrid: bool):
    """``key_fn`` of the network's ranking (BF16 inference, as the budget runs it); with
    ``hybrid`` the larger of the rank fractions by logit and by S/N."""
    def key(r: LevelRecord) -> torch.Tensor:
        lg = ranker.logits(r.folds, r.feats, precision="bf16")
        if not hybrid:
            return lg
        n = lg.numel()
        order = torch.argsort(lg, stable=True)
        rk = torch.empty(n, dtype=torch.float32, device=lg.device)
        rk[order] = torch.arange(n, dtype=torch.float32, device=lg.device) / max(n - 1, 1)
        return torch.maximum(rk, r.feats[:, 3])
    return key


def training_set(recs: list[LevelRecord], labels: list[torch.Tensor],
                 neg_per_level: int = 2048, seed: int = 0, pos_per_level: int = 0):
    """Every positive survivor and up to ``neg_per_level`` random negatives per level;
    with ``pos_per_level`` the positives of a level are repeated up to about that many
    (the first levels hold one or two ancestors each, the last ones hundreds)."""
    dev = recs[0].folds.device
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    fo, fe, la = [], [], []
    for r, lab in zip(recs, labels):
        pos = torch.nonzero(lab > 0).reshape(-1)
        neg = torch.nonzero(lab == 0).reshape(-1)
        if len(neg) > neg_per_level:
            neg = neg[torch.randperm(len(neg), device=dev, generator=g)[:neg_per_level]]
        if 0 < len(pos) < pos_per_level:
            pos = pos.repeat((pos_per_level + len(pos) - 1) // len(pos))
        idx = torch.cat([pos, neg])
        fo.append(r.folds[idx])
        fe.append(r.feats[idx])
        la.append(lab[idx])
    return torch.cat(fo), torch.cat(fe), torch.cat(la)


def train(ranker: BranchRanker, folds: torch.Tensor, feats: torch.Tensor,
          labels: torch.Tensor, epochs: int = 16, batch: int = 4096, lr: float = 1e-3,
          seed: int = 0) -> list[float]:
    """Mini-batch Adam on the class-balanced loss (positives weighted by the ratio of
    negatives to positives); returns the mean loss of every epoch."""
    n = int(folds.shape[0])
    npos = int((labels > 0).sum().item())
    w_pos = float(n - npos) / max(npos, 1)
    g = torch.Generator(device=folds.device)
    g.manual_seed(seed)
    hist = []
    for _ in range(epochs):
        perm = torch.randperm(n, device=folds.device, generator=g)
        tot, nb = 0.0, 0
        for b0 in range(0, n, batch):
            idx = perm[b0:b0 + batch]
            tot += ranker.train_step(folds[idx], feats[idx], labels[idx], lr=lr, w_pos=w_pos)
            nb += 1
        hist.append(tot / max(nb, 1))
    return hist


class SurvivorBudget:
    """Keep, at every level, the ``budget`` survivors of highest key: the ranking
    network's logit when ``ranker`` is given, else the S/N (``pyl_ep_budget_apply``; ties
    at the cut keep the lowest rows). ``budget`` is one number for every level, or a
    schedule: one budget per search level 0..nsegments - 1 (``BudgetSchedule``).
    Installed as ``Pruning.level_hook``: the sync-free level loop applies it on the device
    (``device_budget``), the per-level path calls it."""

    def __init__(self, budget, ranker: BranchRanker | None = None,
                 precision: str = "bf16", hybrid: bool = False,
                 net_until: int | None = None) -> None:
        sched = None
        if np.ndim(budget) > 0:
            sched = np.ascontiguousarray(budget, dtype=np.int64)
            budget = int(sched.max()) if sched.size else 0
        if budget < 1 or (sched is not None and int(sched.min()) < 1):
            msg = "budget must be >= 1"
            raise ValueError(msg)
        if precision not in ("fp32", "bf16"):
            msg = f"unknown precision {precision!r}"
            raise ValueError(msg)
        self.schedule = sched  # per search level, or None
        self.budget = int(budget)  # the largest budget of any level
        self.ranker = ranker
        self.precision = precision  # of the network's forward pass (bf16: MFMA)
        # hybrid: rank by the larger of the rank fractions by logit and by S/N (keeps what
        # either ranking puts near the top)
        if hybrid and ranker is None:
            msg = "a hybrid budget needs the ranking network"
            raise ValueError(msg)
        self.hybrid = bool(hybrid)
        # With the network: the last search level ranked by the network's key; later levels
        # rank by S/N (no inference there). None: every level.
        if net_until is not None and ranker is None:
            msg = "net_until needs the ranking network"
            raise ValueError(msg)
        self.net_until = None if net_until is None else int(net_until)
        self._level_kind = None
        self._cap = -1
        self._s = None
        # Level loop: levels that get the ranking and the cut (None: all), see device_budget.
        self.level_mask: np.ndarray | None = None
        self._mask = None

    def budget_at(self, level: int) -> int:
        """The budget of search level ``level``."""
        if self.schedule is None:
            return self.budget
        return int(self.schedule[min(int(level), len(self.schedule) - 1)])

    def _buffers(self, pr, cap: int):
        """The C struct with scratch for ``cap`` rows (kept alive here)."""
        from pyloki_amd.prune import EpBudget

        cfg = pr.dyp.cfg
        P = cfg.prune_poly_order
        if self._s is None or self._cap < cap or self._s.nbins != cfg.nbins:
            dev = torch.device("cuda")
            t = {"keys": torch.empty(cap, dtype=torch.float32, device=dev),
                 "feats": torch.empty((cap, NFEAT), dtype=torch.float32, device=dev),
                 "kept": torch.empty(cap, dtype=torch.int64, device=dev),
                 "bt_tmp": torch.empty((cap, P + 2), dtype=torch.int32, device=dev),
                 "sel": torch.empty(_BUDGET_SEL_WORDS, dtype=torch.int32, device=dev),
                 "stats": torch.empty(_BUDGET_STATS, dtype=torch.float64, device=dev),
                 "shist": torch.empty(2 * _BUDGET_HIST_BINS, dtype=torch.int32, device=dev),
                 "part": torch.empty(_EP_SCAN_BLOCKS, dtype=torch.int64, device=dev)}
            s = EpBudget()
            s.kind = (_BUDGET_HYBRID if self.hybrid else _BUDGET_CNN) if self.ranker is not None \
                else _BUDGET_SCORE
            s.nbins = cfg.nbins
            s.complex_folds = int(bool(cfg.use_fourier))
            s.budget = self.budget
            s.precision = 1 if self.precision == "bf16" else 0
            s.params = ptr(self.ranker.params) if self.ranker is not None else None
            for name, ten in t.items():
                setattr(s, name, ptr(ten))
            self._t, self._s, self._cap = t, s, cap
        self._s.nlevels = pr.dyp.nsegments - 1
        if self.schedule is not None:
            if len(self.schedule) != self._s.nlevels + 1:
                msg = (f"the budget schedule needs {self._s.nlevels + 1} levels, "
                       f"got {len(self.schedule)}")
                raise ValueError(msg)
            self._s.level_budget = self.schedule.ctypes.data
        else:
            self._s.level_budget = None
        if self.net_until is not None:
            kinds = np.full(self._s.nlevels + 1, _BUDGET_SCORE, dtype=np.int32)
            kinds[:self.net_until + 1] = self._s.kind
            self._level_kind = kinds
            self._s.level_kind = kinds.ctypes.data
        else:
            self._s.level_kind = None
        return self._s

    def device_budget(self, pr, cap: int, level0: int):
        """The budget for the level loop, whose first level is ``level0``. With
        ``level_mask`` set (uint8, one flag per search level 0..nsegments - 1), the loop
        enqueues the ranking and the cut only on flagged levels; another level keeps its
        survivors when they fit the budget and otherwise runs per level, so the result does
        not depend on the mask."""
        s = self._buffers(pr, cap)
        s.level0 = int(level0)
        mask = self.level_mask
        if mask is not None:
            mask = np.ascontiguousarray(mask, dtype=np.uint8)
            if mask.shape != (s.nlevels + 1,):
                raise ValueError(f"level_mask needs {s.nlevels + 1} flags, got {mask.shape}")
            self._mask = mask  # kept alive while the struct points at it
            s.level_mask = mask.ctypes.data
        else:
            s.level_mask = None
        return s

    def __call__(self, pr, tree):
        from pyloki_amd.prune import DeviceTree

        n = tree.size
        if n == 0:
            return tree
        lib = _lib.load()
        s = self._buffers(pr, n)
        P = int(tree.leaves.shape[1]) - 2
        k = min(n, self.budget_at(pr.prune_level))
        dev = tree.leaves.device
        rec = torch.empty((2, _REC_WORDS), dtype=torch.int64, device=dev)
        out = (torch.empty((k, P + 2, 2), dtype=torch.float64, device=dev),
               torch.empty((k, *tree.folds.shape[1:]), dtype=tree.folds.dtype, device=dev),
               torch.empty(k, dtype=torch.float32, device=dev),
               torch.empty((k, P + 2), dtype=torch.int32, device=dev))
        _lib.check(lib.pyl_ep_budget_apply(
            ctypes.byref(s), P, int(pr.prune_level), ptr(tree.leaves), ptr(tree.folds),
            ptr(tree.scores), ptr(tree.backtracks.contiguous()), n, ptr(rec), *map(ptr, out),
            stream_ptr()))
        kept = int(rec[0, _REC_NKEPT].item())
        return DeviceTree(*(o[:kept] for o in out))


def budget_hook(ranker: BranchRanker, budget: int, precision: str = "bf16") -> SurvivorBudget:
    """Keep the ``budget`` survivors the network ranks highest (in their original order)."""
    return SurvivorBudget(budget, ranker, precision)


def score_budget_hook(budget: int) -> SurvivorBudget:
    """The baseline: keep the ``budget`` survivors of highest S/N."""
    return SurvivorBudget(budget)


def _final_keys(leaves: np.ndarray) -> set[bytes]:
    return {r.tobytes() for r in np.ascontiguousarray(leaves)}


def _timed_run(pr, ref_seg: int, hook) -> dict:
    pr.level_hook = hook
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pr.initialize(ref_seg)
    pr.execute_levels()
    torch.cuda.synchronize()
    return {"time_s": time.perf_counter() - t0,
            "trials": int(sum(s.n_leaves for s in pr.pstats[1:])),
            "final": _final_keys(pr.world_tree.leaves.cpu().numpy())}


def compare_searches(pr_factory, ref_seg: int, hooks: dict) -> dict:
    """The exhaustive run and one run per named level hook of the same search. For each
    hook: recall of the exhaustive final candidates, leaves scored (trials), wall time,
    and the ratios to the exhaustive run."""
    ex = _timed_run(pr_factory(), ref_seg, None)
    out = {"exhaustive": {"trials": ex["trials"], "time_s": ex["time_s"],
                          "n_final": len(ex["final"])}}
    for name, hook in hooks.items():
        r = _timed_run(pr_factory(), ref_seg, hook)
        common = len(ex["final"] & r["final"])
        recall = common / len(ex["final"]) if ex["final"] else 1.0
        precision = common / len(r["final"]) if r["final"] else 1.0
        out[name] = {"recall": recall, "precision": precision, "trials": r["trials"],
                     "time_s": r["time_s"],
                     "n_final": len(r["final"]),
                     "trial_ratio": ex["trials"] / max(r["trials"], 1),
                     "speedup": ex["time_s"] / r["time_s"]}
    return out


def train_ranker(make_dyp, seeds, thresholds: np.ndarray, max_sugg: int, epochs: int = 16,
                 batch_size: int = 1024, nbins: int = 64, seed: int = 0,
                 pos_per_level: int = 0) -> tuple[BranchRanker, dict]:
    """The network trained on the exhaustive runs of ``make_dyp(s)`` for ``seeds`` (the
    method's own folds: every level's survivors, labelled by whether they are ancestors of
    a final candidate)."""
    from pyloki_amd.prune import Pruning

    t0 = time.perf_counter()
    parts = []
    for s in seeds:
        dyp = make_dyp(s)
        pr = Pruning(dyp, thresholds, max_sugg=max_sugg, batch_size=batch_size)
        recs = record_levels(pr, dyp.nsegments // 2)
        if recs:
            parts.append(training_set(recs, ancestor_labels(recs), seed=seed + s,
                                      pos_per_level=pos_per_level))
        del recs, pr
    folds = torch.cat([p[0] for p in parts])
    feats = torch.cat([p[1] for p in parts])
    labels = torch.cat([p[2] for p in parts])
    t_data = time.perf_counter() - t0
    ranker = BranchRanker(nbins, seed=seed)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = train(ranker, folds, feats, labels, epochs=epochs, seed=seed)
    torch.cuda.synchronize()
    info = {"samples": int(folds.shape[0]), "positives": int(labels.sum().item()),
            "epochs": epochs, "loss_first": hist[0], "loss_last": hist[-1],
            "data_s": t_data, "train_s": time.perf_counter() - t0, "seeds": list(seeds)}
    return ranker, info


def calibrate_schedules(make_dyp, seeds, thresholds: np.ndarray, max_sugg: int,
                        ranker: BranchRanker | None = None, margin: float = 1.5,
                        floor: int = 512, window: int = 2, batch_size: int = 1024,
                        loss: float = 0.0) -> dict:
    """Budget schedules of the S/N key (and, with ``ranker``, of the network's and the
    hybrid key) from the exhaustive runs of calibration series, which the network has not
    been trained on: ``budget_schedule`` of their ``level_needs`` at ``loss``."""
    from pyloki_amd.prune import Pruning

    keys = {"snr": snr_key}
    if ranker is not None:
        keys["cnn"] = network_keys(ranker, hybrid=False)
        keys["hybrid"] = network_keys(ranker, hybrid=True)
    needs = {k: [] for k in keys}
    survivors = []
    nlevels = 0
    for s in seeds:
        dyp = make_dyp(s)
        nlevels = dyp.nsegments - 1
        pr = Pruning(dyp, thresholds, max_sugg=max_sugg, batch_size=batch_size)
        recs = record_levels(pr, dyp.nsegments // 2)
        weights = descendant_counts(recs)
        for k, fn in keys.items():
            needs[k].append(level_needs(recs, weights, fn, loss))
        survivors.append({r.level: int(r.feats.shape[0]) for r in recs})
        del recs, weights, pr
    out = {k: budget_schedule(v, nlevels, margin, floor, window) for k, v in needs.items()}
    out["needs"] = needs
    out["survivors"] = survivors
    out["nlevels"] = nlevels
    return out


def select_schedule(needs: list[dict], survivors: list[dict], nlevels: int,
                    margins=(1.0, 1.25, 1.5, 2.0), floors=(512, 2048, 4096, 8192, 12288),
                    windows=(0, 1, 2, 3)) -> dict:
    """The ``budget_schedule`` recipe (margin, floor, window) chosen on the calibration
    runs alone: the most calibration series kept when each is left out of the schedule it
    is checked against (a series is kept when the schedule is at least its need at every
    level, which is sufficient: cuts at earlier levels only raise the rank of a kept
    ancestor), and among those the fewest survivors kept (sum over levels and series of
    min(budget, exhaustive survivors)). Returns the schedule from all the series and the
    recipe's figures."""
    n = len(needs)
    best = None
    for m in margins:
        for f in floors:
            for w in windows:
                kept = sum(
                    all(v <= sched[lv] for lv, v in needs[x].items())
                    for x in range(n)
                    for sched in [budget_schedule([needs[y] for y in range(n) if y != x],
                                                  nlevels, m, f, w)])
                sched = budget_schedule(needs, nlevels, m, f, w)
                cost = sum(min(int(sched[lv]), c) for sv in survivors for lv, c in sv.items())
                key = (-kept, cost)
                if best is None or key < best[0]:
                    best = (key, {"margin": m, "floor": f, "window": w, "loo_kept": kept,
                                  "loo_of": n, "kept_survivors": cost, "schedule": sched})
    return best[1]


def budget_experiment(make_dyp, train_seeds, test_seed: int, thresholds: np.ndarray,
                      max_sugg: int, budgets, epochs: int = 16, batch_size: int = 1024,
                      nbins: int = 64, seed: int = 0) -> dict:
    """Train the network on exhaustive runs of ``make_dyp(s)`` for the training seeds,
    then compare the exhaustive run of the test seed with the learned and the S/N budget
    at every budget. The reference segment is the middle one."""
    from pyloki_amd.prune import Pruning

    t0 = time.perf_counter()
    parts = []
    for s in train_seeds:
        dyp = make_dyp(s)
        pr = Pruning(dyp, thresholds, max_sugg=max_sugg, batch_size=batch_size)
        recs = record_levels(pr, dyp.nsegments // 2)
        if recs:
            parts.append(training_set(recs, ancestor_labels(recs), seed=seed + s))
    folds = torch.cat([p[0] for p in parts])
    feats = torch.cat([p[1] for p in parts])
    labels = torch.cat([p[2] for p in parts])
    t_data = time.perf_counter() - t0
    ranker = BranchRanker(nbins, seed=seed)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = train(ranker, folds, feats, labels, epochs=epochs, seed=seed)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    dyp = make_dyp(test_seed)
    hooks = {}
    for b in budgets:
        hooks[f"cnn_{b}"] = budget_hook(ranker, b)
        hooks[f"hybrid_{b}"] = SurvivorBudget(b, ranker, hybrid=True)
        hooks[f"snr_{b}"] = score_budget_hook(b)
    res = compare_searches(lambda: Pruning(dyp, thresholds, max_sugg=max_sugg,
                                           batch_size=batch_size), dyp.nsegments // 2, hooks)
    res["ranker"] = ranker
    res["train"] = {"samples": int(folds.shape[0]), "positives": int(labels.sum().item()),
                    "epochs": epochs, "loss_first": hist[0], "loss_last": hist[-1],
                    "data_s": t_data, "train_s": t_train}
    return res


def batched_throughput(dyps: list, thresholds: np.ndarray, max_sugg: int, hooks: dict,
                       width: int = 4, batch_size: int = 1024) -> dict:
    """Searches per second of each named hook factory (None: exhaustive) over the series
    ``dyps`` of one configuration, ``width`` searches at a time on their own streams
    (``graph.SearchBatch``), and the recall of each run's candidates against the
    exhaustive ones, series by series."""
    from pyloki_amd.graph import SearchBatch

    ref = dyps[0].nsegments // 2
    out, finals = {}, {}
    for name, factory in hooks.items():
        sb = SearchBatch(dyps[0], thresholds, ref, width=width, hook_factory=factory,
                         max_sugg=max_sugg, batch_size=batch_size)
        sb.run(dyps[:width])  # warm-up (captures the graphs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = sb.run(dyps)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        finals[name] = [{r.tobytes() for r in np.ascontiguousarray(x["param_sets"])}
                        for x in res]
        out[name] = {"ms_per_search": dt / len(dyps) * 1e3,
                     "trials": int(sum(sum(st.n_leaves for st in x["level_stats"][1:])
                                       for x in res))}
        del sb
        torch.cuda.empty_cache()
    return _survey_summary(out, finals)


def forest_throughput(dyps: list, thresholds: np.ndarray, max_sugg: int, hooks: dict,
                      batch_size: int = 1024, ntree: int | None = None,
                      capture: bool = False) -> dict:
    """``batched_throughput`` with the series searched ``ntree`` at a time (default: all)
    as the trees of one forest (``forest.ForestSearch``): every launch of a level serves
    all of them. Untimed passes over the series first: one sizes the buffers and, with
    ``capture``, a second captures the level loop as a HIP graph, which the timed pass
    replays."""
    from pyloki_amd.forest import ForestSearch

    ref = dyps[0].nsegments // 2
    T = len(dyps) if ntree is None else int(ntree)
    out, finals = {}, {}
    for name, factory in hooks.items():
        fs = ForestSearch(dyps[0], thresholds, ref, T, hook_factory=factory,
                          capture=capture, max_sugg=max_sugg, batch_size=batch_size)
        for _ in range(2 if capture else 1):  # warm-up (and capture)
            for b0 in range(0, len(dyps), T):
                fs.run(dyps[b0:b0 + T])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = []
        for b0 in range(0, len(dyps), T):
            res += fs.run(dyps[b0:b0 + T])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        finals[name] = [{r.tobytes() for r in np.ascontiguousarray(x["param_sets"])}
                        for x in res]
        out[name] = {"ms_per_search": dt / len(dyps) * 1e3,
                     "trials": int(sum(sum(st.n_leaves for st in x["level_stats"][1:])
                                       for x in res)),
                     "resumes": {int(k): v for k, v in fs.stats["resumes"].items()},
                     "fallbacks": fs.stats["fallbacks"]}
        del fs
        torch.cuda.empty_cache()
    return _survey_summary(out, finals)


def _survey_summary(out: dict, finals: dict) -> dict:
    """Recall of every run's candidates against the exhaustive ones (series by series),
    and the speedup and trial ratio over the exhaustive search."""
    ex = finals.get("exhaustive")
    for name in out:
        if ex is None:
            continue
        hit = sum(len(a & b) for a, b in zip(ex, finals[name]))
        tot = sum(len(a) for a in ex)
        out[name]["recall"] = hit / tot if tot else 1.0
        if name != "exhaustive":
            out[name]["speedup"] = out["exhaustive"]["ms_per_search"] / out[name]["ms_per_search"]
            out[name]["trial_ratio"] = out["exhaustive"]["trials"] / max(out[name]["trials"], 1)
    return out
