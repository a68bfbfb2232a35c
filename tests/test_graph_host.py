"""Host logic of a captured search (no GPU): the per-child arrays fitted to the eager
run's largest level, and the levels that keep a survivor budget's ranking and cut."""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np

from pyloki_amd.cnn import SurvivorBudget
from pyloki_amd.graph import CapturedSearch
from pyloki_amd.prune import PruneStats, ST_OK, ST_OVERFLOW


def _search(pr) -> CapturedSearch:
    cs = object.__new__(CapturedSearch)  # the fitting steps only, no capture
    cs.pr = pr
    return cs


class _Buffers:
    def __init__(self, cap: int) -> None:
        self.cap_child = cap
        self.calls = []

    def set_children(self, cap: int) -> None:
        self.calls.append(cap)
        self.cap_child = cap


def _pr(n_leaves, loop_levels=None, hook=None, cap=2**23):
    stats = [PruneStats(level=0, seg_idx=0, threshold=0.0, score_min=0.0, score_max=0.0,
                        n_branches=1, n_leaves=1, n_leaves_phy=1, n_leaves_surv=1)]
    stats += [PruneStats(level=i + 1, seg_idx=i, threshold=0.0, score_min=0.0, score_max=0.0,
                         n_branches=1, n_leaves=n, n_leaves_phy=n, n_leaves_surv=n)
              for i, n in enumerate(n_leaves)]
    return SimpleNamespace(_pstats=stats, _loop=_Buffers(cap), level_hook=hook,
                           loop_levels=loop_levels or {},
                           dyp=SimpleNamespace(nsegments=len(n_leaves) + 1))


def test_fit_children_twice_the_largest_level():
    pr = _pr([10, 300_000, 70_000])
    _search(pr)._fit_children()
    assert pr._loop.calls == [2**20]  # 2 * 300 000 rounded up to a power of two


def test_fit_children_floor_and_no_growth():
    pr = _pr([5, 7])
    _search(pr)._fit_children()
    assert pr._loop.calls == [2**16]
    pr = _pr([600_000], cap=2**20)
    _search(pr)._fit_children()
    assert pr._loop.calls == []  # never grows the arrays (a larger level hands off)


def test_fit_budget_flags_levels_near_or_above_the_budget():
    hook = SurvivorBudget(1000)
    levels = {1: (ST_OK, 100), 2: (ST_OK, 251), 3: (ST_OK, 250), 4: (ST_OVERFLOW, 0)}
    pr = _pr([0] * 5, loop_levels=levels, hook=hook)
    _search(pr)._fit_budget()
    # level 2 > budget / 4; 4 stopped; 5 has no record (run per level): all flagged
    np.testing.assert_array_equal(hook.level_mask, [0, 0, 1, 0, 1, 1])


def test_fit_budget_keeps_a_given_mask_and_ignores_no_hook():
    hook = SurvivorBudget(10)
    hook.level_mask = np.ones(3, dtype=np.uint8)
    pr = _pr([1, 1], loop_levels={1: (ST_OK, 1), 2: (ST_OK, 1)}, hook=hook)
    _search(pr)._fit_budget()
    np.testing.assert_array_equal(hook.level_mask, [1, 1, 1])
    _search(_pr([1]))._fit_budget()  # no hook: nothing to do
