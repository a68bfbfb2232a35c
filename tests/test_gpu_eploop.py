"""The sync-free EP level loop (``pyl_ep_levels``, ``Pruning(device_loop=True)``) against
the per-level path, which the oracle pins bit for bit (test_gpu_prune.py,
test_gpu_cheby.py, test_gpu_complex.py): reports, scores, per-level statistics and the
backtrack history must be identical, also when the loop stops at a level and hands it to
the host (survivor overflow, child buffer, history buffer, branch error)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import prune_oracle as po
from pyloki_amd.prune import ST_BUDGET, ST_CHILD_CAP, ST_HIST_FULL, ST_OVERFLOW, Pruning
from tests.test_gpu_cheby import _setup
from tests.test_gpu_circ import _setup as _circ_setup

pytestmark = pytest.mark.gpu


def _case(msg: str) -> None:
    print(f"[case] {msg}", flush=True)


def _run(dyp, thresholds, ref, device_loop, caps=None, ascend_levels=None, **kw):
    pr = Pruning(dyp, thresholds, device_loop=device_loop, **kw)
    pr.loop_caps = caps
    res = pr.execute(ref, ascend_levels=ascend_levels)
    return pr, res


def _same(a, b) -> None:
    (pa, ra), (pb, rb) = a, b
    np.testing.assert_array_equal(ra["param_sets"], rb["param_sets"])
    np.testing.assert_array_equal(ra["scores"], rb["scores"])
    np.testing.assert_array_equal(ra["scores_ep"], rb["scores_ep"])
    assert ra["level_stats"] == rb["level_stats"]
    assert pa.prune_level == pb.prune_level and pa.prune_complete == pb.prune_complete
    assert len(pa.backtrack_hist) == len(pb.backtrack_hist)
    for lv, (x, y) in enumerate(zip(pa.backtrack_hist, pb.backtrack_hist)):
        np.testing.assert_array_equal(x.cpu().numpy(), y.cpu().numpy(), err_msg=f"level {lv}")
    for name in ("leaves", "folds", "scores", "backtracks"):
        np.testing.assert_array_equal(getattr(pa.world_tree, name).cpu().numpy(),
                                      getattr(pb.world_tree, name).cpu().numpy(), err_msg=name)


@pytest.mark.parametrize("P,nseg,thr,max_sugg,ref,fourier,basis,caps", [
    (2, 8, 2.5, 2**20, 3, False, "taylor", None),
    (2, 16, 2.0, 64, 7, False, "taylor", None),              # overflow: host replays levels
    (3, 8, 3.0, 2**20, 0, False, "taylor", None),            # jerk, reference at the edge
    (2, 16, 2.0, 2**12, 7, False, "taylor", {"cap_child": 64, "cap_hist": 1}),  # small buffers
    (2, 16, 2.0, 2**20, 7, True, "taylor", None),            # complex folds
    (2, 8, 2.5, 2**20, 3, False, "chebyshev", None),
    (2, 16, 2.0, 64, 7, False, "chebyshev", None),
    (2, 16, 2.0, 2**20, 7, True, "chebyshev", None),
])
def test_loop_equals_per_level_path(gpu_device, P, nseg, thr, max_sugg, ref, fourier, basis,
                                    caps):
    _case(f"ep loop P={P} nseg={nseg} max_sugg={max_sugg} ref={ref} fourier={fourier} "
          f"basis={basis} caps={caps}")
    cfg, dyp, pcfg = _setup(P, 2**17 if nseg <= 8 else 2**18, nseg, use_fourier=fourier)
    thresholds = np.linspace(thr, thr + 3.0, dyp.nsegments - 1)
    kw = {"max_sugg": max_sugg, "batch_size": 64, "poly_basis": basis}
    slow = _run(dyp, thresholds, ref, False, **kw)
    fast = _run(dyp, thresholds, ref, True, caps=caps, **kw)
    _same(fast, slow)
    hand = fast[0].loop_handoffs
    if max_sugg == 64:
        assert hand.get(ST_OVERFLOW, 0) >= 1
    st = fast[0].pstats
    if caps is not None:
        assert hand.get(ST_CHILD_CAP, 0) >= 1
        if sum(s.n_leaves_surv for s in st[1:]) > 4096:  # more than the first history holds
            assert hand.get(ST_HIST_FULL, 0) >= 1
            assert fast[0]._loop.hist.shape[0] > 4096
    print(f"[result] levels {fast[0].prune_level}, leaves {sum(s.n_leaves for s in st)}, "
          f"final {len(fast[1]['scores'])}", flush=True)
    assert len(fast[1]["scores"]) > 0 or max_sugg == 64


def test_loop_matches_oracle_with_ascend_levels(gpu_device):
    _case("ep loop with intermediate ascends vs the oracle")
    cfg, dyp, pcfg = _setup(2, 2**17, 8)
    thresholds = np.linspace(2.0, 5.0, 7)
    fast = _run(dyp, thresholds, 3, True, max_sugg=2**18, ascend_levels=[3, 5])
    slow = _run(dyp, thresholds, 3, False, max_sugg=2**18, ascend_levels=[3, 5])
    _same(fast, slow)
    res = _run(dyp, thresholds, 3, True, max_sugg=2**18)[1]
    exp = po.prune_run(dyp.fold, dyp.param_arr, dyp.dparams_actual, dyp.tseg, thresholds, 3,
                       pcfg, 2**18)
    np.testing.assert_array_equal(res["param_sets"], exp["report"])
    np.testing.assert_array_equal(res["scores"], exp["tree"].scores)


def test_loop_branch_error(gpu_device):
    _case("ep loop branch_max error")
    cfg, dyp, pcfg = _setup(2, 2**17, 8)
    pr = Pruning(dyp, np.full(7, 2.0), max_sugg=2**16)
    pr.initialize(3)
    with torch.no_grad():
        pr.world_tree.leaves[:, 0, 1] = 1e9  # absurd accel cell -> too many children
    with pytest.raises(ValueError, match="increase branch_max"):
        pr._run_levels(7)


@pytest.mark.parametrize("batch_size,max_sugg,fourier,caps", [
    (1024, 2**16, False, None),
    (256, 2**16, False, None),
    (256, 300, False, None),                                  # overflow
    (64, 2**12, False, {"cap_child": 256, "cap_hist": 1}),    # small buffers
    (256, 2**16, True, None),                                 # complex folds
])
def test_circ_loop_equals_per_level_path(gpu_device, batch_size, max_sugg, fourier, caps):
    _case(f"circular ep loop batch_size={batch_size} max_sugg={max_sugg} fourier={fourier} "
          f"caps={caps}")
    cfg, dyp, pcfg = _circ_setup(use_fourier=fourier)
    nseg = dyp.nsegments
    thresholds = np.linspace(2.0, 8.0, nseg - 1)
    kw = {"max_sugg": max_sugg, "batch_size": batch_size}
    slow = _run(dyp, thresholds, nseg // 2, False, **kw)
    fast = _run(dyp, thresholds, nseg // 2, True, caps=caps, **kw)
    _same(fast, slow)
    hand = fast[0].loop_handoffs
    if max_sugg == 300:
        assert hand.get(ST_OVERFLOW, 0) >= 1
    if caps is not None:
        assert hand.get(ST_CHILD_CAP, 0) >= 1
    st = fast[0].pstats
    print(f"[result] levels {fast[0].prune_level}, leaves {sum(s.n_leaves for s in st)}, "
          f"validated {sum(s.n_leaves_phy for s in st)}, final {len(fast[1]['scores'])}, "
          f"hand-offs {hand}", flush=True)


def test_circ_loop_branch_error(gpu_device):
    _case("circular ep loop branch_max error")
    cfg, dyp, pcfg = _circ_setup()
    pr = Pruning(dyp, np.full(dyp.nsegments - 1, 2.0), max_sugg=2**16)
    pr.initialize(dyp.nsegments // 2)
    with torch.no_grad():
        pr.world_tree.leaves[:, 2, 1] = 1e12  # absurd jerk cell -> too many children
    with pytest.raises(ValueError, match="increase branch_max"):
        pr.execute_levels()


# ---- survivor budget -------------------------------------------------------------------

def _tree(n: int, P: int, B: int, seed: int, ties: bool):
    from pyloki_amd.prune import DeviceTree

    rng = np.random.default_rng(seed)
    sc = rng.integers(0, 7, n).astype(np.float32) if ties else rng.normal(size=n).astype(np.float32)
    lv = rng.normal(size=(n, P + 2, 2))
    fo = rng.normal(size=(n, 2, B)).astype(np.float32)
    bt = rng.integers(0, 1000, (n, P + 2)).astype(np.int32)
    t = DeviceTree(*(torch.from_numpy(a).cuda() for a in (lv, fo, sc, bt)))
    return t, sc


class _Pr:  # what a budget reads from its Pruning
    def __init__(self, P, B, nlev=10, level=3):
        class _C:
            prune_poly_order = P
            nbins = B
            use_fourier = False

        class _D:
            cfg = _C()
            nsegments = nlev + 1

        self.dyp = _D()
        self.prune_level = level


@pytest.mark.parametrize("n,budget,ties", [(5000, 700, True), (5000, 700, False),
                                           (300, 700, False), (1, 1, False), (70000, 4096, True)])
def test_score_budget_cut(gpu_device, n, budget, ties):
    from pyloki_amd.cnn import SurvivorBudget

    _case(f"budget cut n={n} budget={budget} ties={ties}")
    P, B = 3, 32
    tree, sc = _tree(n, P, B, n + budget, ties)
    got = SurvivorBudget(budget)(_Pr(P, B), tree)
    order = np.lexsort((np.arange(n), -sc.astype(np.float64)))  # key desc, then row asc
    keep = np.sort(order[:min(n, budget)])
    np.testing.assert_array_equal(got.scores.cpu().numpy(), sc[keep])
    np.testing.assert_array_equal(got.leaves.cpu().numpy(), tree.leaves.cpu().numpy()[keep])
    np.testing.assert_array_equal(got.folds.cpu().numpy(), tree.folds.cpu().numpy()[keep])
    np.testing.assert_array_equal(got.backtracks.cpu().numpy(),
                                  tree.backtracks.cpu().numpy()[keep])


def test_cnn_features_definition(gpu_device):
    from pyloki_amd.cnn import leaf_features

    _case("cnn features")
    P, B, n = 5, 32, 3001
    tree, sc = _tree(n, P, B, 5, False)
    f = leaf_features(tree.leaves, tree.scores, 7, 20).cpu().numpy()
    x = sc.astype(np.float64)
    lo, hi = x.min(), x.max()
    b = np.clip(((x - lo) / (hi - lo) * 1024).astype(np.int64), 0, 1023)
    cnt = np.bincount(b, minlength=1024)
    below = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    rank = (below[b] + 0.5 * (cnt[b] - 1)) / (n - 1)
    z = (x - x.mean()) / (x.std(ddof=1) + 1e-6)
    w = np.log10(1 + np.abs(tree.leaves.cpu().numpy()[:, 1:5, 1])) / 10
    np.testing.assert_allclose(f[:, 0], sc / 10, rtol=1e-6)
    np.testing.assert_allclose(f[:, 1], 7 / 20, rtol=1e-6)
    np.testing.assert_allclose(f[:, 2], z, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(f[:, 3], rank, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(f[:, 4:8], w, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("kind,circular,budget", [("snr", False, 64), ("cnn", False, 64),
                                                  ("hybrid", False, 64), ("hybrid", True, 512),
                                                  ("snr", True, 512), ("cnn", True, 512)])
def test_loop_budget_equals_per_level_path(gpu_device, kind, circular, budget):
    from pyloki_amd.cnn import BranchRanker, SurvivorBudget

    _case(f"ep loop budget kind={kind} circular={circular} budget={budget}")
    if circular:
        cfg, dyp, pcfg = _circ_setup(use_fourier=True)
        ref = dyp.nsegments // 2
        thresholds = np.linspace(2.0, 8.0, dyp.nsegments - 1)
    else:
        cfg, dyp, pcfg = _setup(2, 2**18, 16)
        ref = 7
        thresholds = np.linspace(2.0, 5.0, dyp.nsegments - 1)
    ranker = BranchRanker(cfg.nbins, seed=3) if kind != "snr" else None
    runs = []
    for dl in (False, True):
        pr = Pruning(dyp, thresholds, max_sugg=2**16, batch_size=256, device_loop=dl)
        pr.level_hook = SurvivorBudget(budget, ranker, hybrid=kind == "hybrid")
        runs.append((pr, pr.execute(ref)))
    _same(runs[1], runs[0])
    st = runs[1][0].pstats
    assert max(s.n_leaves_surv for s in st[1:]) <= budget
    assert any(s.n_leaves_surv == budget for s in st[1:])  # the cut was applied
    print(f"[result] levels {runs[1][0].prune_level}, leaves {sum(s.n_leaves for s in st)}, "
          f"final {len(runs[1][1]['scores'])}", flush=True)


@pytest.mark.parametrize("kind,circular,budget", [("snr", False, 64), ("hybrid", True, 512)])
@pytest.mark.parametrize("mask", ["none_flagged", "second_half"])
def test_loop_budget_level_mask(gpu_device, kind, circular, budget, mask):
    """Levels enqueued without the ranking and the cut keep their survivors when they fit
    the budget and otherwise stop and run per level: the result equals the per-level path
    whatever the mask."""
    from pyloki_amd.cnn import BranchRanker, SurvivorBudget

    _case(f"ep loop budget mask={mask} kind={kind} circular={circular} budget={budget}")
    if circular:
        cfg, dyp, pcfg = _circ_setup(use_fourier=True)
        ref = dyp.nsegments // 2
        thresholds = np.linspace(2.0, 8.0, dyp.nsegments - 1)
    else:
        cfg, dyp, pcfg = _setup(2, 2**18, 16)
        ref = 7
        thresholds = np.linspace(2.0, 5.0, dyp.nsegments - 1)
    ranker = BranchRanker(cfg.nbins, seed=3) if kind != "snr" else None
    nlev = dyp.nsegments - 1
    flags = np.zeros(nlev + 1, dtype=np.uint8)
    if mask == "second_half":
        flags[nlev // 2:] = 1
    runs = []
    for dl in (False, True):
        pr = Pruning(dyp, thresholds, max_sugg=2**16, batch_size=256, device_loop=dl)
        pr.level_hook = SurvivorBudget(budget, ranker, hybrid=kind == "hybrid")
        if dl:
            pr.level_hook.level_mask = flags
        runs.append((pr, pr.execute(ref)))
    _same(runs[1], runs[0])
    pr = runs[1][0]
    assert any(s.n_leaves_surv == budget for s in pr.pstats[1:])  # the cut was applied
    if mask == "none_flagged":  # every cut came from a level stopped and run per level
        assert pr.loop_handoffs.get(ST_BUDGET, 0) > 0
    print(f"[result] mask {mask}: handoffs {pr.loop_handoffs}", flush=True)
