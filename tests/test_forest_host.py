"""Host rule of the forest level loop's tree buffers (``ForestSearch._out_buffers``): which
buffer holds each level's kept rows (no GPU). The C loop (``circ_levels``) follows the same
rule; the GPU forest tests check the two against each other through the results."""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np

from pyloki_amd.forest import ForestSearch
from pyloki_amd.prune import BUDGET_SCORE

NET = 3  # a network (hybrid) budget's kind


def _outs(budget, levels, first_out, prev, level_kind=None):
    me = SimpleNamespace(hook=SimpleNamespace(_level_kind=level_kind))
    return ForestSearch._out_buffers(me, budget, levels, first_out, prev)


def test_no_budget_and_snr_budget_alternate():
    for budget in (None, SimpleNamespace(kind=BUDGET_SCORE)):
        assert _outs(budget, range(4), 0, -1) == [0, 1, 0, 1]
        assert _outs(budget, range(3), 0, 1) == [0, 1, 0]
        assert _outs(budget, range(3), 1, 0) == [1, 0, 1]


def test_network_budget_keeps_into_the_input_buffer():
    net = SimpleNamespace(kind=NET)
    assert _outs(net, range(3), 0, -1) == [1, 1, 1]
    assert _outs(net, range(2), 0, 1) == [1, 1]
    assert _outs(net, range(2), 1, 0) == [0, 0]  # materialise into 1, keep into 0


def test_per_level_keys_switch_the_rule():
    net = SimpleNamespace(kind=NET)
    kinds = np.array([NET, NET, BUDGET_SCORE, BUDGET_SCORE, NET], dtype=np.int32)
    # levels 0, 1 by the network (kept into 1), then S/N alternating from 1, then network
    # with its input in 1
    assert _outs(net, range(5), 0, -1, kinds) == [1, 1, 0, 1, 1]
    # a run resumed at level 3 with its input in buffer 0
    assert _outs(net, range(3, 5), 1, 0, kinds) == [1, 1]
