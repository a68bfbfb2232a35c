"""GPU parity of the EP Taylor search (branch, resolve, shift-add, score, threshold,
transform, backtrack, overflow, ascend, report) against the oracle, level by level.
Everything is compared bit-exactly: leaves (FP64), folds and scores (FP32), backtracks."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import prune_oracle as po
from pyloki_amd.config import ParamLimits, PulsarSearchConfig
from pyloki_amd.ffa import DynamicProgramming
from pyloki_amd.prune import Pruning, prune_dyp_tree
from pyloki_amd.simulation.pulse import PulseSignalConfig

pytestmark = pytest.mark.gpu


def _case(msg: str) -> None:
    print(f"[case] {msg}", flush=True)


def _setup(P: int, nsamps: int, nseg_ffa: int, nbins: int = 32, seed: int = 4):
    tsamp = 64e-6
    if P == 2:
        lim = ParamLimits.from_taylor((195.0, 205.0), accel=(-50.0, 50.0)).limits
        mod = {"acc": 12.0}
    else:
        lim = ParamLimits.from_taylor((495.0, 505.0), accel=(-300.0, 300.0), jerk=(-40.0, 40.0)).limits
        mod = {"acc": 30.0, "jerk": 4.0}
    cfg = PulsarSearchConfig(nsamps=nsamps, tsamp=tsamp, nbins=nbins, eta=1.0, param_limits=lim,
                             bseg_brute=nsamps // 256, bseg_ffa=nsamps // nseg_ffa,
                             prune_poly_order=P, ducy_max=0.3, wtsp=1.5, use_fourier=False,
                             branch_max=32)
    sim = PulseSignalConfig(period=1 / (lim[-1][0] + 5.37), dt=tsamp, nsamps=nsamps, snr=30.0,
                            ducy=0.1, mod_kwargs=mod)
    ts = sim.generate_analytic(seed=seed)
    dyp = DynamicProgramming(ts, cfg)
    dyp.initialize()
    dyp.execute()
    pcfg = {"poly_order": P, "nbins": nbins, "eta": cfg.eta, "branch_max": cfg.branch_max,
            "counts_init": np.array([len(a) for a in dyp.param_arr]),
            "param_limits": cfg.param_limits, "widths": cfg.score_widths,
            "tiling": cfg.tiling_strategy}
    return cfg, dyp, pcfg


def _compare_tree(gtree, otree, what: str) -> None:
    np.testing.assert_array_equal(gtree.scores.cpu().numpy(), otree.scores, err_msg=what)
    np.testing.assert_array_equal(gtree.leaves.cpu().numpy(), otree.leaves, err_msg=what)
    np.testing.assert_array_equal(gtree.folds.cpu().numpy(), otree.folds, err_msg=what)
    np.testing.assert_array_equal(gtree.backtracks.cpu().numpy(), otree.backtracks, err_msg=what)


@pytest.mark.parametrize("P,nsamps,nseg,thr,max_sugg,ref", [
    (2, 2**17, 8, 2.5, 2**20, 3),     # no overflow
    (2, 2**17, 16, 2.0, 300, 7),      # survivor buffer overflows: batch-sequential rule
    (3, 2**17, 8, 3.0, 2**20, 0),     # jerk (Taylor order 3), reference segment at the edge
])
def test_prune_levels_bit_exact(gpu_device, P, nsamps, nseg, thr, max_sugg, ref):
    _case(f"prune P={P} nsamps={nsamps} nseg={nseg} max_sugg={max_sugg} ref={ref}")
    cfg, dyp, pcfg = _setup(P, nsamps, nseg)
    thresholds = np.linspace(thr, thr + 3.0, nseg - 1)
    dyp_fold = dyp.fold
    exp = po.prune_run(dyp_fold, dyp.param_arr, dyp.dparams_actual, dyp.tseg, thresholds, ref,
                       pcfg, max_sugg, batch_size=64)
    pr = Pruning(dyp, thresholds, max_sugg=max_sugg, batch_size=64)
    pr.initialize(ref)
    for level in range(1, nseg):
        pr.execute_iter()
        ot = exp["trees"][level - 1] if level - 1 < len(exp["trees"]) else None
        if ot is None:
            assert pr.prune_complete
            break
        _compare_tree(pr.world_tree, ot, f"level {level}")
        st, ost = pr.pstats[level], exp["stats"][level - 1]
        assert st.n_leaves == ost.n_leaves
        assert np.float32(st.score_min) == np.float32(ost.score_min)
        assert np.float32(st.score_max) == np.float32(ost.score_max)
        if pr.prune_complete:
            break
    pr._do_ascend()
    t = exp["tree"]
    if t.size:
        np.testing.assert_array_equal(pr.world_tree.folds.cpu().numpy(), t.folds)
        np.testing.assert_array_equal(pr.world_tree.scores.cpu().numpy(), t.scores)
        np.testing.assert_array_equal(pr.world_tree.scores_ep.cpu().numpy(), t.scores_ep)


def test_prune_execute_and_report(gpu_device, tmp_path):
    _case("prune execute + report")
    cfg, dyp, pcfg = _setup(2, 2**17, 8)
    thresholds = np.linspace(2.0, 5.0, 7)
    res = Pruning(dyp, thresholds, max_sugg=2**18).execute(3)
    exp = po.prune_run(dyp.fold, dyp.param_arr, dyp.dparams_actual, dyp.tseg, thresholds, 3,
                       pcfg, 2**18)
    np.testing.assert_array_equal(res["param_sets"], exp["report"])
    np.testing.assert_array_equal(res["scores"], exp["tree"].scores)
    best = int(np.argmax(res["scores"]))
    assert abs(res["param_sets"][best, -3, 0] - 200.37) < 0.05
    path = prune_dyp_tree(dyp, thresholds, n_runs=2, max_sugg=2**18, outdir=tmp_path)
    assert path.endswith("test_pruning_nstages_8_results.npz")
    from pyloki_amd.io.cands import read_results
    from pyloki_amd.periodogram import PruningStatsPlotter, ScatteredPeriodogram

    runs = read_results(path)["runs"]
    assert list(runs) == ["002_00", "005_01"]
    r5 = Pruning(dyp, thresholds, max_sugg=2**18).execute(5)
    np.testing.assert_array_equal(runs["005_01"]["param_sets"], r5["param_sets"])
    np.testing.assert_array_equal(runs["005_01"]["scores"], r5["scores"])
    np.testing.assert_array_equal(runs["005_01"]["scores_ep"], r5["scores_ep"])
    sp = ScatteredPeriodogram.load(path)
    assert sp.n_runs == 2 and len(sp.data) == len(runs["002_00"]["scores"]) + len(r5["scores"])
    assert PruningStatsPlotter.load(path).n_runs == 2
    assert (tmp_path / "test_pruning_nstages_8_log.txt").read_text().startswith("Pruning log")


def test_prune_branch_max_error(gpu_device):
    _case("prune branch_max error")
    cfg, dyp, pcfg = _setup(2, 2**17, 8)
    pr = Pruning(dyp, np.full(7, 2.0), max_sugg=2**16)
    pr.initialize(3)
    with torch.no_grad():
        pr.world_tree.leaves[:, 0, 1] = 1e9  # absurd accel cell -> too many children
    with pytest.raises(ValueError, match="increase branch_max"):
        pr.execute_iter()
