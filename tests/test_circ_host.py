"""Host logic of the circular EP driver (no GPU): the final child order of a level."""

from __future__ import annotations

import numpy as np
import torch

from pyloki_amd.prune import circ_child_positions


def _loop_positions(nchild1, expand, cr_num, bs):
    """Direct restatement: per batch, kept held children in order, then the crackle
    children of the expanding ones (circular.py:359-394)."""
    par = np.repeat(np.arange(len(nchild1)), nchild1)
    pos = np.empty(len(par), dtype=np.int64)
    cur = 0
    for b0 in range(0, len(nchild1), bs):
        sel = np.where((par >= b0) & (par < b0 + bs))[0]
        for s in sel:
            if not expand[s]:
                pos[s] = cur
                cur += 1
        for s in sel:
            if expand[s]:
                pos[s] = cur
                cur += cr_num[par[s]]
    return pos, cur


def test_circ_child_positions_match_loop():
    rng = np.random.default_rng(9)
    for trial in range(30):
        n = int(rng.integers(1, 300))
        nchild1 = rng.integers(1, 7, n)
        cr_num = rng.integers(0, 5, n).astype(np.int32)
        par = np.repeat(np.arange(n), nchild1)
        expand = (rng.random(len(par)) < 0.3) & (cr_num[par] > 0)
        bs = int(rng.integers(1, 80)) if trial % 3 else 1024
        exp_pos, exp_total = _loop_positions(nchild1, expand, cr_num, min(bs, n))
        off1 = torch.zeros(n + 1, dtype=torch.int64)
        off1[1:] = torch.from_numpy(np.cumsum(nchild1))
        pos, total, nb = circ_child_positions(off1, torch.from_numpy(par.astype(np.int32)),
                                              torch.from_numpy(expand.astype(np.uint8)),
                                              torch.from_numpy(cr_num), bs)
        assert total == exp_total
        assert nb == -(-n // max(1, min(bs, n)))
        np.testing.assert_array_equal(pos.numpy(), exp_pos)
