"""Many EP searches of one configuration through one level loop (a "forest").

A survey searches many series of one configuration: the same grids, reference segment,
coordinates and thresholds, one FFA fold per series. The reference runs such searches as
separate tasks (``prune.py:892-922``: one ``Pruning`` per process and reference segment).
``ForestSearch`` runs ``ntree`` of them through every launch of the circular level loop
(``pyl_circ_forest_levels``): a level's parents are the trees' parents one tree after the
other, each tree keeps its own batches, survivor limit and (with a survivor budget) its
own ranking and cut, so every tree's result equals ``Pruning(dyp_t, ...).execute(ref_seg)``
bit for bit while the chain of small dependent launches per level is paid once for all
trees.

A level the loop cannot finish on the device is handled as the single-tree loop does:
buffers too small (children, history, the forest's survivor buffer) are grown and the
forest resumes at that level; a level enqueued without the budget's cut whose survivors
exceed the budget is flagged and resumed with the cut; a branch error or a tree whose
passing children exceed its ``max_sugg`` (the reference's overflow rule) or whose level is
too large for the forest's packed per-batch sums sends every tree on through its own
``Pruning`` from that level.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from pyloki_amd import _lib
from pyloki_amd.device import ptr, stream_ptr
from pyloki_amd.prune import (BUDGET_SCORE, EP_COMPLEX, EP_FOREST_MAX, EP_FOREST_TAB, EP_SCAN_BLOCKS,
                              REC_HIST, REC_NCHILD, REC_STATUS, REC_WORDS, ST_BUDGET,
                              ST_CHILD_CAP, ST_HIST_FULL, ST_OK, ST_OVERFLOW, TAB_CB, TAB_KO,
                              TAB_PO, TAB_SO, TAB_VB, BUDGET_HIST_BINS, BUDGET_SEL_WORDS,
                              BUDGET_STATS, DeviceTree, EpForest, LoopBuffers, PruneCoords,
                              Pruning, PruneStats, PruneStatsCollection,
                              poly_taylor_report_batch)
from pyloki_amd.utils.snail import MiddleOutScheme


def _pow2(n: int, floor: int = 2**16) -> int:
    return max(floor, 1 << max(int(n) - 1, 0).bit_length())


class ForestSearch:
    """``Pruning(dyp_t, threshold_scheme, **kw).execute(ref_seg)`` for ``ntree`` series of
    the configuration of ``dyp`` at once (circular basis). ``hook_factory`` builds the
    survivor budget (``cnn.SurvivorBudget``) shared by the trees, or None (exhaustive).
    The buffers are sized from an eager run of ``dyp``; a run that needs more grows them."""

    def __init__(self, dyp, threshold_scheme: np.ndarray, ref_seg: int, ntree: int,
                 hook_factory=None, capture: bool = False, **kw) -> None:
        if not 1 <= ntree <= EP_FOREST_MAX:
            msg = f"ntree must be in [1, {EP_FOREST_MAX}]"
            raise ValueError(msg)
        self.T = int(ntree)
        self.ref_seg = int(ref_seg)
        self.hook = hook_factory() if hook_factory is not None else None
        self.prs = [Pruning(dyp, threshold_scheme, device_loop=True, **kw)
                    for _ in range(self.T)]
        pr0 = self.prs[0]
        if not pr0._circular:
            msg = "the forest runs the circular basis (prune_poly_order 5)"
            raise ValueError(msg)
        for pr in self.prs:
            pr.level_hook = self.hook
        self.nlast = dyp.nsegments - 1
        self._nseg = dyp.nsegments
        # Sizes from an eager run of the configuration.
        pr0.execute(self.ref_seg)
        self._fit_mask(pr0)
        st = pr0.pstats
        n_seed = int(st[0].n_leaves)
        surv = [int(v[1]) for v in pr0.loop_levels.values()] + [
            int(s.n_leaves_surv) for s in st[1:]]
        child = max([int(s.n_leaves) for s in st[1:]] + [1])
        hist = sum(int(s.n_leaves_surv) for s in st[1:])
        T = self.T
        self.max_tot = min(T * pr0.max_sugg, _pow2(2 * T * max(surv + [1])))
        self.cap_par = max(self.max_tot, T * n_seed)
        self.cap_child = _pow2(2 * T * child)
        self.cap_hist = self.max_tot + _pow2(2 * T * hist)
        self._buf = None
        self._alloc()
        # Stacked folds of the trees, and the seed rows (same leaves for every tree).
        self.fold = torch.empty((T, *dyp.fold_device.shape), dtype=dyp.fold_device.dtype,
                                device=dyp.fold_device.device)
        pr0.initialize(self.ref_seg)
        seed = pr0.world_tree
        self.n_seed = seed.size
        self.seed_leaves = seed.leaves.repeat(T, 1, 1).contiguous()
        self.seed_folds = torch.empty((T * self.n_seed, *seed.folds.shape[1:]),
                                      dtype=seed.folds.dtype, device=seed.folds.device)
        self.seed_scores = torch.empty(T * self.n_seed, dtype=torch.float32,
                                       device=seed.folds.device)
        self.stats = {"resumes": {}, "fallbacks": 0}
        # capture: a run's first enqueue of the level loop (every level from the seed) is
        # captured as a HIP graph once the buffers have settled, and replayed by later runs
        # of as many trees; resumes run eagerly (see _first_enqueue).
        self.capture = bool(capture)
        self._graph = None
        self._graph_key = None
        self._graph_run = None
        self._warm_key = None

    # -- buffers ---------------------------------------------------------------------
    def _fit_mask(self, pr) -> None:
        """Budget flags as ``graph.CapturedSearch`` sets them: the cut only on the levels
        whose survivors (before the cut) exceeded a quarter of the budget, or where the
        eager run stopped (results do not depend on the flags)."""
        hook = self.hook
        if hook is None or not hasattr(hook, "level_mask") or hook.level_mask is not None:
            return
        mask = np.zeros(self.nlast + 1, dtype=np.uint8)
        for level in range(1, self.nlast + 1):
            st = pr.loop_levels.get(level)
            mask[level] = st is None or st[0] != 0 or st[1] > hook.budget_at(level) // 4
        hook.level_mask = mask

    def _alloc(self) -> None:
        pr0 = self.prs[0]
        dev = torch.device("cuda")
        T = self.T
        self._buf = LoopBuffers(pr0._P, pr0._new_folds, self.cap_par, self.cap_child,
                                self.cap_hist, self.max_tot, self.nlast, False, True)
        f = EpForest()
        f.ntree = T
        f.max_sugg = pr0.max_sugg
        self._t = {
            "tab": torch.empty((self.nlast + 1, EP_FOREST_TAB, T + 1), dtype=torch.int64,
                               device=dev),
            "minmax": torch.empty((self.nlast, T, 2), dtype=torch.float32, device=dev),
            "sel": torch.empty((T, BUDGET_SEL_WORDS), dtype=torch.int32, device=dev),
            "stats": torch.empty((T, BUDGET_STATS), dtype=torch.float64, device=dev),
            "shist": torch.empty((T, 2 * BUDGET_HIST_BINS), dtype=torch.int32, device=dev),
            "part": torch.empty((T, EP_SCAN_BLOCKS), dtype=torch.int64, device=dev),
            "kcount": torch.empty(T, dtype=torch.int64, device=dev),
        }
        for name, t in self._t.items():
            setattr(f, name, ptr(t))
        self._f = f

    # -- one run -----------------------------------------------------------------------
    def run(self, dyps: list) -> list[dict]:
        """The searches of ``dyps`` (at most ``ntree`` series of the configuration): the
        result dicts of ``Pruning.execute``, in order."""
        if not 1 <= len(dyps) <= self.T:
            msg = f"a forest of {self.T} trees takes 1..{self.T} series"
            raise ValueError(msg)
        T = len(dyps)
        for t, d in enumerate(dyps):
            if d.fold_device.shape != self.fold.shape[1:] or d.fold_device.dtype != self.fold.dtype:
                msg = "the series' fold does not match the forest's configuration"
                raise ValueError(msg)
            self.fold[t].copy_(d.fold_device)
        prs = self.prs[:T]
        self._buf.hist_next = 0  # the previous run's backtracks are no longer needed
        self._seed(prs, dyps)
        n = self.n_seed
        state = {"l0": 1, "leaves": self.seed_leaves[:T * n], "folds": self.seed_folds[:T * n],
                 "sizes": [n] * T, "buf": -1}
        self._last = None
        self._fell_back = False
        run = self._first_enqueue(prs, state)
        while True:
            state = self._finish(prs, run)
            if state is None:
                break
            run = self._enqueue(prs, state)
        return self._ascend_report(prs)

    def _capture_key(self, T: int) -> tuple:
        """Everything a captured first enqueue depends on: the trees, the loop's and the
        forest's device buffers, and the budget's buffers, constants and host arrays."""
        # Fields that _enqueue sets from the tree count and the hook, both in the key.
        skip = {"ntree", "tree_stride", "budget", "timer"}

        def vals(st) -> tuple:
            out = []
            for name, _ in st._fields_:
                if name in skip:
                    continue
                v = getattr(st, name)
                out.append(tuple(v) if isinstance(v, ctypes.Array) else v)
            return tuple(out)

        key = [T, vals(self._buf.s), vals(self._buf.c), vals(self._f)]
        hook = self.hook
        s = getattr(hook, "_s", None) if hook is not None else None
        if s is not None:
            key += [getattr(s, f) for f in ("kind", "nbins", "complex_folds", "nlevels",
                                            "precision", "budget", "params", "keys", "feats",
                                            "kept", "bt_tmp", "sel", "stats", "shist", "part")]
            for arr in (getattr(hook, "schedule", None), getattr(hook, "_level_kind", None),
                        getattr(hook, "level_mask", None)):
                key.append(None if arr is None else bytes(np.ascontiguousarray(arr)))
        return tuple(key)

    def _first_enqueue(self, prs, state: dict) -> dict:
        """The run's first enqueue: eager, or (capture) the replay of the graph captured
        for the same key. A key is run eagerly once before it is captured (that run sizes
        the buffers and allocates the budget's scratch); the capture is replayed at once."""
        if not self.capture:
            return self._enqueue(prs, state)
        key = self._capture_key(len(prs))
        if self._graph is not None and key == self._graph_key:
            self._graph.replay()
            return self._graph_run
        self._graph = None
        if key != self._warm_key:
            self._warm_key = key
            return self._enqueue(prs, state)
        g = torch.cuda.CUDAGraph()
        torch.cuda.synchronize()
        with torch.cuda.graph(g):
            run = self._enqueue(prs, state)
        self._graph, self._graph_key, self._graph_run = g, key, run
        g.replay()
        return run

    def _seed(self, prs, dyps) -> None:
        """Every tree's seed: the reference segment's rows of its fold, scored in one
        launch; each Pruning's host state as initialize() leaves it."""
        from pyloki_amd import kernels

        T, n = len(prs), self.n_seed
        pr0 = prs[0]
        ref_idx = MiddleOutScheme(self._nseg, self.ref_seg, dyps[0].tseg).ref_idx
        folds = self.seed_folds[:T * n]
        folds.view(T, n, -1).copy_(self.fold[:T, ref_idx].reshape(T, n, -1))
        scores = self.seed_scores[:T * n]
        if pr0._complex:
            kernels.score_boxcar_complex(folds, pr0._widths, pr0.dyp.cfg.nbins,
                                         mode=_lib.PYL_SCORE_MAX, out=scores)
        else:
            kernels.score_boxcar(folds, pr0._widths, mode=_lib.PYL_SCORE_MAX, out=scores)
        mm = torch.stack([scores.view(T, n).amin(1), scores.view(T, n).amax(1)], 1).cpu()
        bt = torch.zeros((n, pr0._P + 2), dtype=torch.int32, device=folds.device)
        for t, (pr, d) in enumerate(zip(prs, dyps)):
            pr._dyp = d
            pr._scheme = MiddleOutScheme(d.nsegments, self.ref_seg, d.tseg)
            pr._prune_complete = False
            pr._prune_level = 0
            pr._backtrack_hist = []
            sl = slice(t * n, (t + 1) * n)
            pr._world_tree = DeviceTree(self.seed_leaves[sl], folds[sl], scores[sl], bt)
            pr._pstats = PruneStatsCollection()
            pr._pstats.append(PruneStats(level=0, seg_idx=pr.scheme.ref_idx, threshold=0,
                                         score_min=float(mm[t, 0]), score_max=float(mm[t, 1]),
                                         n_branches=n, n_leaves=n, n_leaves_phy=n,
                                         n_leaves_surv=n))
            pr.loop_levels = {}
            pr.loop_handoffs = {}

    def _enqueue(self, prs, state: dict) -> dict:
        lib = _lib.load()
        pr0 = prs[0]
        T = len(prs)
        buf = self._buf
        l0 = state["l0"]
        levels = range(l0, self.nlast + 1)
        nlev = len(levels)
        fold = self.fold
        per_tree = int(fold[0].numel())
        seg_floats = per_tree // self._nseg
        if fold.is_complex():
            per_tree *= 2
            seg_floats *= 2
        coords = (PruneCoords * nlev)(*[pr0._coords(lv) for lv in levels])
        thr = np.ascontiguousarray(pr0.threshold_scheme[l0 - 1:], dtype=np.float64)
        segs = np.array([pr0.scheme.get_segment_idx(lv) for lv in levels], dtype=np.int64)
        sizes = np.ascontiguousarray(state["sizes"], dtype=np.int64)
        budget = None
        if self.hook is not None:
            budget = self.hook.device_budget(pr0, buf.cap_par, l0)
        buf.s.budget = ctypes.addressof(budget) if budget is not None else None
        first_out = 1 - state["buf"] if state["buf"] >= 0 else 0
        outs = self._out_buffers(budget, levels, first_out, state["buf"])
        self._f.ntree = T
        self._f.tree_stride = per_tree
        _lib.check(lib.pyl_circ_forest_levels(
            ctypes.byref(pr0._pp), ctypes.byref(pr0._qq), EP_COMPLEX if pr0._complex else 0,
            coords, thr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            segs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), nlev, ptr(fold), seg_floats,
            ptr(state["leaves"]), ptr(state["folds"]),
            sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), buf.hist_next, first_out,
            ctypes.byref(buf.s), ctypes.byref(buf.c), ctypes.byref(self._f), stream_ptr()))
        return {"levels": levels, "thr": thr, "segs": segs, "budget": budget,
                "first_out": first_out, "outs": outs, "state": state}

    def _finish(self, prs, run: dict) -> dict | None:
        """Advance every tree over the run's finished levels (one transfer); the state to
        resume the forest from, or None when done."""
        buf = self._buf
        T = len(prs)
        levels, thr, segs = run["levels"], run["thr"], run["segs"]
        budget, first_out = run["budget"], run["first_out"]
        nlev = len(levels)
        t_ = self._t
        tab_w = EP_FOREST_TAB * (T + 1)
        info = torch.cat([
            buf.rec[:nlev + 1].reshape(-1),
            t_["tab"].view(-1)[:(nlev + 1) * tab_w],
            t_["minmax"].view(torch.int64).view(-1)[:nlev * T],
        ]).cpu().numpy()
        rec = info[:(nlev + 1) * REC_WORDS].reshape(nlev + 1, REC_WORDS)
        o1 = (nlev + 1) * REC_WORDS
        tab = info[o1:o1 + (nlev + 1) * tab_w].reshape(nlev + 1, EP_FOREST_TAB, T + 1)
        mm = info[o1 + (nlev + 1) * tab_w:].view(np.float32).reshape(nlev, T, 2)
        status = rec[:nlev, REC_STATUS]
        bad = np.flatnonzero(status != ST_OK)
        i_stop = int(bad[0]) if len(bad) else nlev
        stop = int(status[i_stop]) if len(bad) else ST_OK
        self._advance(prs, run, rec, tab, mm, i_stop)
        if all(pr._prune_complete for pr in prs):
            return None
        if stop == ST_OK:
            return None
        level = levels[i_stop]
        self.stats["resumes"][stop] = self.stats["resumes"].get(stop, 0) + 1
        # The stopped level's input: the run's input, or the previous level's output.
        po = tab[i_stop][TAB_PO]
        sizes = [int(po[t + 1] - po[t]) for t in range(T)]
        if i_stop == 0:
            st = dict(run["state"])
            leaves, folds, src = st["leaves"], st["folds"], st["buf"]
        else:
            src = run["outs"][i_stop - 1]
            leaves, folds = buf.leaves[src], buf.folds[src]
        state = {"l0": level, "leaves": leaves, "folds": folds, "sizes": sizes, "buf": src}
        if stop == ST_CHILD_CAP and self.cap_child < 2**30:
            nc = int(rec[i_stop][REC_NCHILD])
            self.cap_child = _pow2(2 * max(self.cap_child, nc))
            buf.set_children(self.cap_child)
            return state
        if stop == ST_HIST_FULL:
            self.cap_hist *= 2
            buf.new_hist(self.cap_hist)
            return state
        limit = T * self.prs[0].max_sugg  # no tree over its limit fits in less
        if stop == ST_OVERFLOW and self.max_tot < limit:  # the forest's survivor buffer
            ntot = sum(sizes)
            keep = (leaves[:ntot].clone(), folds[:ntot].clone())
            self.max_tot = min(limit, 2 * self.max_tot)
            self.cap_par = max(self.cap_par, self.max_tot)
            self.cap_hist = max(self.cap_hist, 2 * self.max_tot)
            self._alloc()
            state.update(leaves=keep[0], folds=keep[1], buf=-1)
            return state
        if stop == ST_BUDGET and self.hook is not None and self.hook.level_mask is not None:
            self.hook.level_mask[level] = 1
            return state
        # A branch error or a tree over its survivor limit: every tree goes on alone.
        self.stats["fallbacks"] += 1
        self._fell_back = True
        self._fallback(prs)
        return None

    def _advance(self, prs, run: dict, rec, tab, mm, nfin: int) -> None:
        """Every tree over the run's first nfin levels, as ``Pruning._end_level`` would:
        its level statistics and backtrack rows; its world tree (a view of the level's
        output rows) after its last level."""
        if nfin == 0:
            return
        buf = self._buf
        T = len(prs)
        levels, thr, segs = run["levels"], run["thr"], run["segs"]
        outs = run["outs"]
        d = np.diff(tab[:nfin], axis=2)  # [level][table][tree]: per-tree counts
        k0 = tab[:nfin, TAB_KO, :T]
        hist0 = rec[:nfin, REC_HIST]
        mmn = mm[:nfin].astype(np.float64)
        for t, pr in enumerate(prs):
            if pr._prune_complete:
                continue
            bts = pr._backtrack_hist
            stats = pr._pstats
            last = nfin - 1
            for i in range(nfin):
                m = int(d[i, TAB_KO, t])
                nv = int(d[i, TAB_VB, t])
                stats.append(PruneStats(
                    level=levels[i], seg_idx=int(segs[i]), threshold=float(thr[i]),
                    score_min=float(mmn[i, t, 0]) if nv else 0.0,
                    score_max=float(mmn[i, t, 1]) if nv else 0.0,
                    n_branches=int(d[i, TAB_PO, t]), n_leaves=int(d[i, TAB_CB, t]),
                    n_leaves_phy=nv, n_leaves_surv=m))
                pr.loop_levels[levels[i]] = (ST_OK, int(d[i, TAB_SO, t]))
                if m == 0:
                    last = i
                    break
                a = int(hist0[i] + k0[i, t])
                bts.append(buf.hist.narrow(0, a, m))
            pr._prune_level = levels[last]
            o = outs[last]
            a, m = int(k0[last, t]), int(d[last, TAB_KO, t])
            sc = buf.scores[o].narrow(0, a, m)
            pr._world_tree = DeviceTree(buf.leaves[o].narrow(0, a, m),
                                        buf.folds[o].narrow(0, a, m), sc,
                                        buf.hist.narrow(0, int(hist0[last]) + a, m),
                                        scores_ep=sc)
            pr._prune_complete = m == 0
        buf.hist_next = int(hist0[nfin - 1] + tab[nfin - 1, TAB_KO, T])
        self._last = (outs[nfin - 1], tab[nfin - 1, TAB_KO].copy())

    def _out_buffers(self, budget, levels, first_out: int, prev: int) -> list[int]:
        """The tree buffer that holds each level's kept rows, by the level loop's rule: a
        level without a budget, or one ranked by S/N (cut before materialising), writes the
        buffer its input is not in (the first level: first_out); a level ranked by a
        network materialises into the buffer its input is not in and gathers the kept rows
        into the other (the input's, or 1 for an outside input). ``prev``: the buffer of
        the run's input, or -1."""
        lk = getattr(self.hook, "_level_kind", None) if budget is not None else None
        outs = []
        for lv in levels:
            kind = None if budget is None else (int(lk[lv]) if lk is not None else budget.kind)
            if budget is None or kind == BUDGET_SCORE:
                out = first_out if prev < 0 else 1 - prev
            else:
                out = 0 if prev == 0 else 1
            outs.append(out)
            prev = out
        return outs

    def _ascend_report(self, prs) -> list[dict]:
        """The ascend of every tree in one launch, and the reports with one transfer, when
        the trees left are the forest's last level's rows; else tree by tree."""
        pr0 = prs[0]
        T = len(prs)
        live = [t for t, pr in enumerate(prs) if not pr._prune_complete]
        last = getattr(self, "_last", None)
        if self._fell_back or not live or last is None or any(
                prs[t]._prune_level != self.nlast for t in live):
            out = []
            for pr in prs:
                pr._do_ascend()
                out.append(pr.report(self.ref_seg))
            return out
        o, ko = last
        ko = np.ascontiguousarray(ko, dtype=np.int64)
        n = int(ko[T])
        buf = self._buf
        leaves = buf.leaves[o][:n]
        scores_ep = buf.scores[o][:n].clone()
        folds = pr0._new_folds(n, leaves.device)
        scores = torch.empty(n, dtype=torch.float32, device=leaves.device)
        sch = prs[live[0]].scheme
        seg_idx, seg_coords = sch.get_segment_coords_so_far(self.nlast)
        t0_cur, _ = sch.get_previous_coord(self.nlast + 1, pr0._use_moving_grid)
        idx = np.ascontiguousarray(seg_idx, dtype=np.int64)
        t0 = np.ascontiguousarray(seg_coords[:, 0], dtype=np.float64)
        per_tree = int(self.fold[0].numel()) * (2 if self.fold.is_complex() else 1)
        lib = _lib.load()
        _lib.check(lib.pyl_circ_forest_ascend(
            ctypes.byref(pr0._pp), ctypes.byref(pr0._qq), EP_COMPLEX if pr0._complex else 0,
            ptr(leaves), ko.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), T, ptr(self.fold),
            per_tree, pr0._ncoords, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            t0.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(idx), t0_cur, ptr(folds),
            ptr(scores), stream_ptr()))
        host = torch.cat([leaves.reshape(-1).view(torch.int32),
                          torch.stack([scores, scores_ep]).reshape(-1).view(torch.int32)]
                         ).cpu().numpy()  # one transfer
        nl = 2 * int(np.prod(leaves.shape[1:]))  # 32-bit words per leaf row
        lv = host[:n * nl].view(np.float64).reshape(n, *leaves.shape[1:])
        sc2 = host[n * nl:].view(np.float32).reshape(2, n)
        rep = poly_taylor_report_batch(lv) if n else lv
        out = []
        for t, pr in enumerate(prs):
            a, b = int(ko[t]), int(ko[t + 1])
            tree = pr._world_tree
            if not pr._prune_complete:
                tree.scores_ep = scores_ep[a:b]
                tree.folds = folds[a:b]
                tree.scores = scores[a:b]
            out.append({"run_name": f"{self.ref_seg:03d}_00", "scheme": pr.scheme.data.copy(),
                        "param_sets": rep[a:b], "scores": sc2[0, a:b],
                        "scores_ep": sc2[1, a:b], "level_stats": pr._pstats})
        return out

    def _fallback(self, prs) -> None:
        """Every unfinished tree through its own Pruning from the stopped level, one tree
        after the other in one set of single-tree loop buffers. A tree's world tree holds
        its parents of that level (views of the forest's buffers)."""
        shared = self.prs[0]._loop
        for pr in prs:
            if pr._prune_complete:
                continue
            w = pr._world_tree
            pr._world_tree = DeviceTree(w.leaves.clone(), w.folds.clone(), w.scores.clone(),
                                        w.backtracks, scores_ep=w.scores_ep.clone())
            if shared is not None:
                pr._loop = shared
            pr.execute_levels()
            shared = pr._loop
            w = pr._world_tree
            pr._world_tree = DeviceTree(w.leaves.clone(), w.folds.clone(), w.scores.clone(),
                                        w.backtracks.clone(), scores_ep=w.scores_ep.clone())
