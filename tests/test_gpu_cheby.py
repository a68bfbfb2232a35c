"""GPU parity of the Chebyshev-basis EP search against the oracle: every level's leaves,
folds, scores and backtracks bit-exactly (real folds), the ascend, and the report; with
complex folds, each level's oracle input is the GPU's previous tree and scores and folds
agree within FP32 tolerance."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import cheby_oracle as co
from oracle import prune_oracle as po
from pyloki_amd.config import ParamLimits, PulsarSearchConfig
from pyloki_amd.ffa import DynamicProgramming
from pyloki_amd.prune import Pruning
from tests.test_gpu_complex import _EP_CAP, _align, _cap, _close, _host_tree

pytestmark = pytest.mark.gpu


def _case(msg: str) -> None:
    print(f"[case] {msg}", flush=True)


def _setup(P: int, nsamps: int, nseg_ffa: int, use_fourier: bool = False, seed: int = 4,
           acc_max: float = 50.0):
    from pyloki_amd.simulation.pulse import PulseSignalConfig

    tsamp = 64e-6
    if P == 2:
        lim = ParamLimits.from_taylor((195.0, 205.0), accel=(-acc_max, acc_max)).limits
        mod = {"acc": 12.0}
    else:
        lim = ParamLimits.from_taylor((495.0, 505.0), accel=(-300.0, 300.0),
                                      jerk=(-40.0, 40.0)).limits
        mod = {"acc": 30.0, "jerk": 4.0}
    cfg = PulsarSearchConfig(nsamps=nsamps, tsamp=tsamp, nbins=32, eta=1.0, param_limits=lim,
                             bseg_brute=nsamps // 256, bseg_ffa=nsamps // nseg_ffa,
                             prune_poly_order=P, ducy_max=0.3, wtsp=1.5,
                             use_fourier=use_fourier, branch_max=32)
    sim = PulseSignalConfig(period=1 / (lim[-1][0] + 5.37), dt=tsamp, nsamps=nsamps, snr=30.0,
                            ducy=0.1, mod_kwargs=mod)
    dyp = DynamicProgramming(sim.generate_analytic(seed=seed), cfg)
    dyp.initialize()
    dyp.execute()
    pcfg = {"poly_order": P, "nbins": cfg.nbins, "eta": cfg.eta, "branch_max": cfg.branch_max,
            "counts_init": np.array([len(a) for a in dyp.param_arr]),
            "param_limits": cfg.param_limits, "widths": cfg.score_widths,
            "tiling": cfg.tiling_strategy}
    return cfg, dyp, pcfg


def _compare_tree(gtree, otree, what: str) -> None:
    np.testing.assert_array_equal(gtree.scores.cpu().numpy(), otree.scores, err_msg=what)
    np.testing.assert_array_equal(gtree.leaves.cpu().numpy(), otree.leaves, err_msg=what)
    np.testing.assert_array_equal(gtree.folds.cpu().numpy(), otree.folds, err_msg=what)
    np.testing.assert_array_equal(gtree.backtracks.cpu().numpy(), otree.backtracks, err_msg=what)


@pytest.mark.parametrize("P,nsamps,nseg,thr,max_sugg,ref,acc_max", [
    (2, 2**17, 8, 2.5, 2**20, 3, 50.0),     # no overflow
    (2, 2**20, 32, 1.0, 2**20, 16, 200.0),  # 31 levels, a few thousand children per level
    (2, 2**17, 16, 2.0, 300, 7, 50.0),      # survivor buffer overflows: batch-sequential rule
    (3, 2**17, 8, 3.0, 2**20, 0, 50.0),     # jerk, reference segment at the edge
])
def test_cheby_prune_levels_bit_exact(gpu_device, P, nsamps, nseg, thr, max_sugg, ref, acc_max):
    _case(f"chebyshev prune P={P} nsamps={nsamps} nseg={nseg} max_sugg={max_sugg} ref={ref}")
    cfg, dyp, pcfg = _setup(P, nsamps, nseg, acc_max=acc_max)
    thresholds = np.linspace(thr, thr + 3.0, nseg - 1)
    exp = co.cheby_prune_run(dyp.fold, dyp.param_arr, dyp.dparams_actual, dyp.tseg, thresholds,
                             ref, pcfg, max_sugg, batch_size=64)
    pr = Pruning(dyp, thresholds, max_sugg=max_sugg, batch_size=64, poly_basis="chebyshev")
    pr.initialize(ref)
    np.testing.assert_array_equal(pr.world_tree.leaves.cpu().numpy(), exp["seed"].leaves)
    np.testing.assert_array_equal(pr.world_tree.scores.cpu().numpy(), exp["seed"].scores)
    largest = 0
    for level in range(1, nseg):
        pr.execute_iter()
        ot = exp["trees"][level - 1] if level - 1 < len(exp["trees"]) else None
        if ot is None:
            assert pr.prune_complete
            break
        _compare_tree(pr.world_tree, ot, f"level {level}")
        st, ost = pr.pstats[level], exp["stats"][level - 1]
        assert st.n_leaves == ost.n_leaves
        largest = max(largest, st.n_leaves)
        assert np.float32(st.score_min) == np.float32(ost.score_min)
        assert np.float32(st.score_max) == np.float32(ost.score_max)
        if pr.prune_complete:
            break
    print(f"[result] levels {pr.prune_level}, largest level {largest} children", flush=True)
    pr._do_ascend()
    t = exp["tree"]
    if t.size:
        np.testing.assert_array_equal(pr.world_tree.folds.cpu().numpy(), t.folds)
        np.testing.assert_array_equal(pr.world_tree.scores.cpu().numpy(), t.scores)
        np.testing.assert_array_equal(pr.world_tree.scores_ep.cpu().numpy(), t.scores_ep)


def test_cheby_execute_report(gpu_device):
    _case("chebyshev execute + report")
    cfg, dyp, pcfg = _setup(2, 2**17, 8)
    thresholds = np.linspace(2.0, 5.0, 7)
    res = Pruning(dyp, thresholds, max_sugg=2**18, poly_basis="chebyshev").execute(3)
    exp = co.cheby_prune_run(dyp.fold, dyp.param_arr, dyp.dparams_actual, dyp.tseg, thresholds,
                             3, pcfg, 2**18)
    np.testing.assert_array_equal(res["param_sets"], exp["report"])
    np.testing.assert_array_equal(res["scores"], exp["tree"].scores)
    best = int(np.argmax(res["scores"]))
    assert abs(res["param_sets"][best, -3, 0] - 200.37) < 0.05
    tay = Pruning(dyp, thresholds, max_sugg=2**18).execute(3)
    bt = int(np.argmax(tay["scores"]))
    assert abs(res["param_sets"][best, -3, 0] - tay["param_sets"][bt, -3, 0]) < 0.01


def test_cheby_complex_levels(gpu_device):
    """Complex folds: each level's oracle input is the GPU's previous tree (capped)."""
    _case("chebyshev prune, complex folds")
    cfg, dyp, pcfg = _setup(2, 2**18, 16, use_fourier=True)
    nseg = dyp.nsegments
    thresholds = np.linspace(2.0, 5.0, nseg - 1)
    pr = Pruning(dyp, thresholds, max_sugg=2**20, batch_size=256, poly_basis="chebyshev")
    pr.initialize(7)
    scheme = po.MiddleOut(nseg, 7, dyp.tseg)
    fold = dyp.fold
    worst, skipped = 0.0, 0
    for level in range(1, nseg):
        _cap(pr)
        tin = _host_tree(pr.world_tree)
        pr.execute_iter()
        thr = float(thresholds[level - 1])
        exp, st, _ = co.cheby_prune_level(tin, fold[scheme.get_segment_idx(level)],
                                          co.level_coords(scheme, level), thr, pcfg, 2**20, 256)
        got = _host_tree(pr.world_tree)
        pairs, sk = _align(got, exp, thr, f"level {level}")
        skipped += sk
        gi, ej = pairs[:, 0], pairs[:, 1]
        np.testing.assert_array_equal(got.leaves[gi], exp.leaves[ej])
        worst = max(worst, _close(got.scores[gi], exp.scores[ej], f"level {level} scores"))
        _close(got.folds[gi], exp.folds[ej], f"level {level} folds")
        assert pr.pstats[level].n_leaves == st.n_leaves
        if pr.prune_complete:
            break
    print(f"[result] complex chebyshev levels {pr.prune_level}: worst score error / max "
          f"{worst:.2e}, one-sided near-threshold survivors {skipped} (cap {_EP_CAP})",
          flush=True)
    assert pr.prune_level > nseg // 2
    _cap(pr)
    tin = _host_tree(pr.world_tree)
    pr._do_ascend()
    lvl = pr.prune_level
    idx_seg, coord_seg = scheme.get_segment_coords_so_far(lvl)
    exp = co.cheby_ascend(tin, fold, idx_seg, coord_seg, scheme.get_coord(lvl), pcfg)
    _close(pr.world_tree.folds.cpu().numpy(), exp.folds, "ascend folds")
    _close(pr.world_tree.scores.cpu().numpy(), exp.scores, "ascend scores")


def test_cheby_branch_max_error(gpu_device):
    _case("chebyshev branch_max error")
    cfg, dyp, pcfg = _setup(2, 2**17, 8)
    pr = Pruning(dyp, np.full(7, 2.0), max_sugg=2**16, poly_basis="chebyshev")
    pr.initialize(3)
    with torch.no_grad():
        pr.world_tree.leaves[:, 0, 1] = 1e12  # absurd alpha_2 cell -> too many children
    with pytest.raises(ValueError, match="increase branch_max"):
        pr.execute_iter()


def test_cheby_invalid_configuration(gpu_device):
    cfg, dyp, pcfg = _setup(2, 2**17, 8)
    with pytest.raises(ValueError, match="Invalid pruning configuration"):
        Pruning(dyp, np.full(7, 2.0), poly_basis="legendre")
