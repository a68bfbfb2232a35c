"""An EP search captured once as a HIP graph and replayed for many series.

A survey searches many series of one configuration: the same grids, segments, reference
segment and thresholds, so every level's constants and launches are the same and only the
FFA fold differs. ``CapturedSearch`` records the seed scoring and every level of the
sync-free level loop (``Pruning._enqueue_levels``) once into a ``torch.cuda.CUDAGraph``;
``run(dyp)`` copies the series' fold into the graph's input buffer, replays the graph
(about 20 kernels per level, enqueued with one call) and reads the level records once.

A level the loop hands back to the host (survivor overflow, buffers too small, a branch
error) is finished eagerly, exactly as ``Pruning.execute`` would, and the graph is
captured again before the next run when the buffers changed. Results equal
``Pruning(dyp, ...).execute(ref_seg)`` bit for bit. No memset node is recorded: the loop
clears its records with a kernel.
"""

from __future__ import annotations

import numpy as np
import torch

from pyloki_amd import _lib
from pyloki_amd.prune import DeviceTree, Pruning, PruneStats, PruneStatsCollection


class _FoldView:
    """A DynamicProgramming whose fold is the graph's input buffer."""

    def __init__(self, dyp, fold: torch.Tensor) -> None:
        self._dyp = dyp
        self.fold_device = fold

    def __getattr__(self, name):
        return getattr(self._dyp, name)


class CapturedSearch:
    """``Pruning(dyp, threshold_scheme, **kw).execute(ref_seg)`` for many series of the
    configuration of ``dyp``, replayed from one captured graph."""

    def __init__(self, dyp, threshold_scheme: np.ndarray, ref_seg: int, level_hook=None,
                 loop_caps: dict | None = None, **kw) -> None:
        self.fold = dyp.fold_device.clone()
        self.pr = Pruning(_FoldView(dyp, self.fold), threshold_scheme, device_loop=True, **kw)
        self.pr.level_hook = level_hook
        self.pr.loop_caps = loop_caps
        self.ref_seg = int(ref_seg)
        self.graph = None
        self.captures = 0
        # An eager run sizes the loop's buffers (and the budget's scratch) for this data.
        self.pr.execute(self.ref_seg)
        if loop_caps is None or "cap_child" not in loop_caps:
            self._fit_children()
        self._fit_budget()

    def _fit_children(self) -> None:
        """Per-child arrays for twice the eager run's largest level (at least 2^16): the
        children and materialize launches cover the capacity, and on small levels the
        blocks beyond the level's children only cost dispatches. A series with a larger
        level hands that level to the host, which grows the arrays; the graph is then
        captured again."""
        need = max([int(st.n_leaves) for st in self.pr._pstats[1:]] + [1])
        cap = max(2**16, 1 << (2 * need - 1).bit_length())
        buf = self.pr._loop
        if buf is not None and cap < buf.cap_child:
            buf.set_children(cap)

    def _fit_budget(self) -> None:
        """A survivor budget's ranking and cut only on the levels where the eager run's
        survivors (before the cut) exceeded a quarter of the budget, or where it stopped:
        the other levels keep their survivors in one launch instead of about 15 (a search's
        survivors rise to the budget and fall well below it on the last levels). A series
        whose survivors exceed the budget on an unflagged level runs that level per level,
        so results do not depend on this."""
        hook = self.pr.level_hook
        if hook is None or not hasattr(hook, "level_mask") or hook.level_mask is not None:
            return
        nlast = self.pr.dyp.nsegments - 1
        mask = np.zeros(nlast + 1, dtype=np.uint8)
        for level in range(1, nlast + 1):
            st = self.pr.loop_levels.get(level)
            mask[level] = st is None or st[0] != 0 or st[1] > hook.budget_at(level) // 4
        hook.level_mask = mask

    def _buffer_ptrs(self):
        """Everything the captured launches hold by address or by value that may change
        between runs: the loop's buffers, and a survivor budget's scratch pointers, budget,
        kind, precision and level flags. The graph is captured again when any changes."""
        buf = self.pr._loop
        key = [int(buf.s.hist or 0), int(buf.s.cap_child), int(buf.s.c_score or 0)]
        hook = self.pr.level_hook
        s = getattr(hook, "_s", None)
        if s is not None:
            key += [int(getattr(s, f) or 0) for f in ("params", "keys", "feats", "kept",
                                                     "bt_tmp", "sel", "stats", "shist",
                                                     "part")]
            key += [int(s.budget), int(s.kind), int(s.precision), int(s.nlevels)]
            sched = getattr(hook, "schedule", None)
            key.append(None if sched is None else bytes(sched))
            lk = getattr(hook, "_level_kind", None)
            key.append(None if lk is None else bytes(lk))
            mask = getattr(hook, "level_mask", None)
            key.append(None if mask is None
                       else bytes(np.ascontiguousarray(mask, dtype=np.uint8)))
        return tuple(key)

    def _capture(self) -> None:
        pr = self.pr
        pr.initialize(self.ref_seg)
        seed = pr.world_tree
        self.seed = DeviceTree(seed.leaves.clone(), seed.folds.clone(), seed.scores.clone(),
                               seed.backtracks.clone())
        self.seed.scores_ep = self.seed.scores  # the replayed seed scores
        n = seed.size
        nlast = pr.dyp.nsegments - 1
        buf = pr._loop_buffers(n, nlast)
        buf.hist_next = 0
        pr.timing = False
        from pyloki_amd import kernels

        rows = self.fold[pr.scheme.ref_idx].reshape(n, 2, -1)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.seed.folds.copy_(rows)
            if pr._complex:
                kernels.score_boxcar_complex(self.seed.folds, pr._widths, pr.dyp.cfg.nbins,
                                             mode=_lib.PYL_SCORE_MAX, out=self.seed.scores)
            else:
                kernels.score_boxcar(self.seed.folds, pr._widths, mode=_lib.PYL_SCORE_MAX,
                                     out=self.seed.scores)
            self.run_ctx = pr._enqueue_levels(1, nlast, self.seed)
        torch.cuda.synchronize()
        self.graph = g
        self.captures += 1
        self.ptrs = self._buffer_ptrs()

    def run(self, dyp) -> dict:
        """The search of ``dyp`` (same configuration): the result dict of
        ``Pruning.execute``. Device views in it (the world tree) live until the next run."""
        self.launch(dyp)
        return self.finish()

    def launch(self, dyp) -> None:
        """Enqueue the search of ``dyp`` on the current stream (no host synchronisation
        once the graph exists)."""
        src = dyp.fold_device
        if src.shape != self.fold.shape or src.dtype != self.fold.dtype:
            msg = "the series' fold does not match the captured configuration"
            raise ValueError(msg)
        if self.graph is None or self._buffer_ptrs() != self.ptrs:
            self._capture()
        torch.cuda.current_stream().wait_stream(torch.cuda.default_stream())
        self.fold.copy_(src)
        self.graph.replay()

    def finish(self) -> dict:
        """Read the launched search's records and complete it (on the current stream)."""
        pr = self.pr
        # The host side of initialize(): the same scheme, the replayed seed.
        pr._prune_complete = False
        pr._prune_level = 0
        pr._backtrack_hist = []
        pr._world_tree = self.seed
        pr._loop.hist_next = 0
        n = self.seed.size
        pr._pstats = PruneStatsCollection()
        pr._pstats.append(PruneStats(level=0, seg_idx=pr.scheme.ref_idx, threshold=0,
                                     score_min=self.seed.score_min,
                                     score_max=self.seed.score_max, n_branches=n, n_leaves=n,
                                     n_leaves_phy=n, n_leaves_surv=n))
        pr._finish_levels(self.run_ctx)
        pr.execute_levels()  # levels after a hand-off, if any
        pr._do_ascend()
        return pr.report(self.ref_seg)


class SearchBatch:
    """Many series of one configuration, ``width`` searches at a time on their own
    streams: each stream replays its own captured search (its own buffers), so the small
    dependent launches of one search's levels overlap those of the others and the GPU's
    CUs stay busy on levels that a single search cannot fill. Results equal
    ``Pruning.execute`` of each series."""

    def __init__(self, dyp, threshold_scheme: np.ndarray, ref_seg: int, width: int = 8,
                 hook_factory=None, loop_caps: dict | None = None, **kw) -> None:
        self.streams = [torch.cuda.Stream() for _ in range(width)]
        self.searches = []
        for st in self.streams:
            with torch.cuda.stream(st):
                hook = hook_factory() if hook_factory is not None else None
                self.searches.append(CapturedSearch(dyp, threshold_scheme, ref_seg,
                                                    level_hook=hook, loop_caps=loop_caps,
                                                    **kw))
        torch.cuda.synchronize()

    def run(self, dyps: list) -> list[dict]:
        out = []
        w = len(self.searches)
        for b0 in range(0, len(dyps), w):
            chunk = dyps[b0:b0 + w]
            for cs, st, d in zip(self.searches, self.streams, chunk):
                with torch.cuda.stream(st):
                    cs.launch(d)
            for cs, st, _ in zip(self.searches, self.streams, chunk):
                with torch.cuda.stream(st):
                    out.append(cs.finish())
        torch.cuda.default_stream().wait_stream(torch.cuda.current_stream())
        for st in self.streams:
            torch.cuda.default_stream().wait_stream(st)
        return out


class CapturedFFABatch:
    """One step of the batched FFA search (``ffa.ffa_batch``: the brute fold of every
    series, every level and the scoring) captured once as a HIP graph.

    The series' samples live in static device buffers (one pair per series, the lengths
    fixed at capture); ``run(series)`` copies new samples of the same lengths into them
    (or ``replay()`` reuses the resident ones), replays the graph and returns the per-width
    S/N grids, which are the graph's own output buffers (overwritten by the next replay).
    The step's launches carry every constant by value (the brute-fold series table, the
    level tables' addresses, the tile), allocate nothing outside the capture's memory pool
    and record no memset node, so a replay equals an eager ``ffa_batch`` bit for bit.
    """

    def __init__(self, plan, series: list[tuple[torch.Tensor, torch.Tensor]],
                 widths: np.ndarray, stdnoise: float = 1.0) -> None:
        from pyloki_amd.ffa import ffa_batch

        self.plan = plan
        self.widths = np.asarray(widths)
        self.stdnoise = float(stdnoise)
        self.inputs = [(e.clone(), v.clone()) for e, v in series]
        self.lengths = [int(e.shape[0]) for e, _ in series]
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):  # the warm-up fills torch's allocator before capture
            ffa_batch(plan, self.inputs, self.widths, self.stdnoise)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.outputs = ffa_batch(plan, self.inputs, self.widths, self.stdnoise)

    def replay(self) -> list[torch.Tensor]:
        self.graph.replay()
        return self.outputs

    def run(self, series: list[tuple[torch.Tensor, torch.Tensor]]) -> list[torch.Tensor]:
        if [int(e.shape[0]) for e, _ in series] != self.lengths:
            msg = "the series must have the lengths the graph was captured with"
            raise ValueError(msg)
        for (de, dv), (e, v) in zip(self.inputs, series):
            de.copy_(e)
            dv.copy_(v)
        return self.replay()
