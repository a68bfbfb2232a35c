"""Many circular-orbit EP searches through one level loop (``forest.ForestSearch``)
against eager ``Pruning.execute`` of each series alone: reports, scores, EP scores, level
statistics and every level's backtracks identical, with and without a survivor budget
(S/N, network, hybrid), on real and complex folds; also when levels stop on the device
(buffers too small: grown and resumed; a tree over its survivor limit: the trees go on
alone) and for a forest that runs fewer series than it holds."""

from __future__ import annotations

import numpy as np
import pytest

from pyloki_amd.forest import ForestSearch
from pyloki_amd.prune import (ST_CHILD_CAP, ST_HIST_FULL, ST_OVERFLOW, Pruning)
from tests.test_gpu_circ import _setup as _circ_setup

pytestmark = pytest.mark.gpu


def _case(msg: str) -> None:
    print(f"[case] {msg}", flush=True)


def _series(seeds, use_fourier: bool):
    from pyloki_amd.ffa import DynamicProgramming
    from pyloki_amd.simulation.pulse import PulseSignalConfig

    cfg, dyp0, _ = _circ_setup(use_fourier=use_fourier)
    out = []
    for s in seeds:
        sim = PulseSignalConfig(period=1 / 200.17, dt=cfg.tsamp, nsamps=cfg.nsamps, snr=30.0,
                                ducy=0.1,
                                mod_type="circular",
                                mod_kwargs={"p_orb": 30.0, "x_orb": 0.002, "psi": 0.1 * s})
        dyp = DynamicProgramming(sim.generate_analytic(seed=s), cfg)
        dyp.initialize()
        dyp.execute()
        out.append(dyp)
    return cfg, out


def _eager(dyp, thresholds, ref, hook_factory, **kw):
    pr = Pruning(dyp, thresholds, **kw)
    pr.level_hook = hook_factory() if hook_factory else None
    res = pr.execute(ref)
    return res, [b.cpu().numpy() for b in pr.backtrack_hist]


def _check(fs: ForestSearch, got: list, dyps, thresholds, ref, hook_factory, **kw) -> None:
    for t, d in enumerate(dyps):
        exp, bt = _eager(d, thresholds, ref, hook_factory, **kw)
        what = f"tree {t}"
        np.testing.assert_array_equal(got[t]["param_sets"], exp["param_sets"], err_msg=what)
        np.testing.assert_array_equal(got[t]["scores"], exp["scores"], err_msg=what)
        np.testing.assert_array_equal(got[t]["scores_ep"], exp["scores_ep"], err_msg=what)
        assert got[t]["level_stats"] == exp["level_stats"], what
        fb = [b.cpu().numpy() for b in fs.prs[t].backtrack_hist]
        assert len(fb) == len(bt), what
        for lv, (a, b) in enumerate(zip(fb, bt)):
            np.testing.assert_array_equal(a, b, err_msg=f"{what} level {lv + 1} backtracks")


def _budget(kind: str, nbins: int):
    from pyloki_amd.cnn import BranchRanker, SurvivorBudget

    if kind == "none":
        return None
    if kind == "snr":
        return lambda: SurvivorBudget(256)
    ranker = BranchRanker(nbins, seed=2)
    if kind == "hybrid_then_snr":  # the network's key on levels 0..3, S/N after
        return lambda: SurvivorBudget(256, ranker, hybrid=True, net_until=3)
    return lambda: SurvivorBudget(256, ranker, hybrid=kind == "hybrid")


@pytest.mark.parametrize("use_fourier", [False, True])
@pytest.mark.parametrize("kind", ["none", "snr", "cnn", "hybrid", "hybrid_then_snr"])
def test_forest_equals_eager(gpu_device, use_fourier, kind):
    _case(f"forest of 4 circular trees, complex={use_fourier}, budget={kind}")
    cfg, dyps = _series((3, 4, 5, 6), use_fourier)
    thresholds = np.linspace(2.0, 8.0, dyps[0].nsegments - 1)
    ref = dyps[0].nsegments // 2
    factory = _budget(kind, cfg.nbins)
    kw = {"max_sugg": 2**20, "batch_size": 256}
    fs = ForestSearch(dyps[0], thresholds, ref, 4, hook_factory=factory, **kw)
    got = fs.run(dyps)
    print(f"[result] resumes {fs.stats}", flush=True)
    _check(fs, got, dyps, thresholds, ref, factory, **kw)
    assert fs.stats["fallbacks"] == 0
    # a second run, fewer series than the forest holds, in another order
    sub = [dyps[2], dyps[0], dyps[3]]
    got = fs.run(sub)
    _check(fs, got, sub, thresholds, ref, factory, **kw)


@pytest.mark.parametrize("kind", ["none", "snr"])
def test_forest_resumes_with_grown_buffers(gpu_device, kind):
    _case(f"forest with small buffers, budget={kind}")
    cfg, dyps = _series((3, 4, 5), True)
    thresholds = np.linspace(2.0, 8.0, dyps[0].nsegments - 1)
    ref = dyps[0].nsegments // 2
    factory = _budget(kind, cfg.nbins)
    kw = {"max_sugg": 2**20, "batch_size": 256}
    fs = ForestSearch(dyps[0], thresholds, ref, 3, hook_factory=factory, **kw)
    fs.max_tot, fs.cap_child, fs.cap_hist = 512, 2048, 1024
    fs.cap_par = max(fs.max_tot, 3 * fs.n_seed)
    fs._alloc()
    got = fs.run(dyps)
    print(f"[result] {fs.stats}", flush=True)
    _check(fs, got, dyps, thresholds, ref, factory, **kw)
    res = fs.stats["resumes"]
    assert res.get(ST_CHILD_CAP) and res.get(ST_OVERFLOW) and res.get(ST_HIST_FULL), res
    assert fs.stats["fallbacks"] == 0


def test_forest_tree_over_its_limit_goes_on_alone(gpu_device):
    _case("forest with max_sugg 64 (the overflow rule)")
    cfg, dyps = _series((3, 4, 5), False)
    thresholds = np.linspace(2.0, 8.0, dyps[0].nsegments - 1)
    ref = dyps[0].nsegments // 2
    kw = {"max_sugg": 64, "batch_size": 32}
    fs = ForestSearch(dyps[0], thresholds, ref, 3, **kw)
    got = fs.run(dyps)
    print(f"[result] {fs.stats}", flush=True)
    assert fs.stats["fallbacks"] == 1
    _check(fs, got, dyps, thresholds, ref, None, **kw)


@pytest.mark.parametrize("use_fourier", [False, True])
def test_forest_net_until_switches_the_key(gpu_device, use_fourier):
    """A network budget that ranks by S/N from level net_until + 1 on: with net_until = -1
    it keeps exactly what the S/N budget keeps (the network never runs), and with net_until
    past the last level exactly what the plain hybrid budget keeps."""
    from pyloki_amd.cnn import BranchRanker, SurvivorBudget

    _case(f"forest net_until limits, complex={use_fourier}")
    cfg, dyps = _series((3, 4, 5), use_fourier)
    thresholds = np.linspace(2.0, 8.0, dyps[0].nsegments - 1)
    ref = dyps[0].nsegments // 2
    kw = {"max_sugg": 2**20, "batch_size": 256}
    ranker = BranchRanker(cfg.nbins, seed=2)
    last = dyps[0].nsegments

    def run(factory):
        fs = ForestSearch(dyps[0], thresholds, ref, 3, hook_factory=factory, **kw)
        return fs.run(dyps)

    for a, b in ((lambda: SurvivorBudget(256, ranker, hybrid=True, net_until=-1),
                  lambda: SurvivorBudget(256)),
                 (lambda: SurvivorBudget(256, ranker, hybrid=True, net_until=last),
                  lambda: SurvivorBudget(256, ranker, hybrid=True))):
        ga, gb = run(a), run(b)
        for t in range(len(dyps)):
            np.testing.assert_array_equal(ga[t]["param_sets"], gb[t]["param_sets"])
            np.testing.assert_array_equal(ga[t]["scores"], gb[t]["scores"])
            assert ga[t]["level_stats"] == gb[t]["level_stats"]


@pytest.mark.parametrize("kind", ["none", "snr", "hybrid"])
def test_captured_forest_replays_equal_eager(gpu_device, kind):
    """A forest whose level loop is captured as a HIP graph: the first run eager, the second
    captured and replayed, the third a replay; each run on other series equals the eager
    per-tree searches bit for bit."""
    _case(f"captured forest replays, budget={kind}")
    cfg, dyps = _series((3, 4, 5, 6), False)
    thresholds = np.linspace(2.0, 8.0, dyps[0].nsegments - 1)
    ref = dyps[0].nsegments // 2
    factory = _budget(kind, cfg.nbins)
    kw = {"max_sugg": 2**20, "batch_size": 256}
    fs = ForestSearch(dyps[0], thresholds, ref, 3, hook_factory=factory, capture=True, **kw)
    for run, sub in enumerate(([dyps[0], dyps[1], dyps[2]], [dyps[3], dyps[1], dyps[0]],
                               [dyps[2], dyps[3], dyps[1]])):
        got = fs.run(sub)
        print(f"[result] run {run}: graph {fs._graph is not None}, {fs.stats}", flush=True)
        assert (fs._graph is not None) == (run >= 1)  # eager, captured, replayed
        _check(fs, got, sub, thresholds, ref, factory, **kw)
