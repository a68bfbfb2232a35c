"""GPU parity at the full sizes of the benchmark workloads.

- FFA acceleration batch (32 seeded series of 2^21..2^23 samples, the benchmark's step):
  sampled series of every length, including the last 2^23 series of the batch (the largest
  byte offsets of the concatenated folds, beyond 2^32), bit-exact against the oracle.
- FFA jerk search of a 2^24-sample series (nbins 128, accel +-30, jerk +-2), FFA capped at
  bseg_ffa = N/8 (10 levels, final grid 8 x 10 x 57553 over 8 segments): the S/N grid of
  segment 0 bit-exact against the oracle in FP32 and in BF16 storage; the error of both
  against the FP64 oracle stated and bounded.
- EP jerk search (2^23 samples, 64 segments, nbins 64): the first 12 levels bit-exact
  against the oracle; every 8th level after them bit-exact with the oracle continuing from
  the GPU's tree before it (cut to its first 8192 rows, on both sides, beside the search);
  the level loop equal
  to the per-level path over all 63 levels.
- EP circular-orbit search (2^24 samples, 128 segments, nbins 64), real folds (bit-exact)
  and complex folds (the reference's default; FP32 tolerance): each of the first 20
  levels and every 16th level after them against the oracle run from the GPU's previous
  tree (deep levels cut to 4096 rows), the level loop against the per-level path, and the
  learned survivor budget's recall of the exhaustive candidates on held-out series.
"""

from __future__ import annotations

import functools

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from oracle import prune_oracle as po

pytestmark = pytest.mark.gpu


def _case(msg: str) -> None:
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module", autouse=True)
def _oracle_threads():
    orc.set_threads(orc.cpu_count())


# ------------------------------------------------------------------ FFA accel batch


def test_ffa_bench_batch_sampled_series_bit_exact(gpu_device):
    import bench
    from pyloki_amd import autotune
    from pyloki_amd.ffa import FFAPlan, ffa_batch

    _case("FFA bench batch: 32 series of 2^21..2^23, sampled series vs oracle")
    series = bench.make_series(32)
    lens = [ts.nsamps for ts in series]
    cfg = bench.make_config(max(lens))
    plan = FFAPlan(cfg)
    dev = [(torch.from_numpy(ts.ts_e).cuda(), torch.from_numpy(ts.ts_v).cuda()) for ts in series]
    widths = cfg.score_widths
    # Batch order is longest first and stable: the last 2^23 series sits at the highest
    # offsets of every level's fold.
    order = sorted(range(len(lens)), key=lambda i: -lens[i])
    big = [i for i in order if lens[i] == 2**23]
    picks = [big[0], big[-1]] + [next(i for i in order if lens[i] == n) for n in (2**22, 2**21)]
    exp = {}
    for i in picks:
        sub = bench.make_config(lens[i])
        exp[i] = orc.ffa_search_snr(series[i].ts_e, series[i].ts_v, sub).reshape(-1, len(widths))
    for rows in (8, 16):
        autotune.set_tile_rows(rows)
        try:
            out = ffa_batch(plan, dev, widths)
        finally:
            autotune.set_tile_rows(0)
        for i in picks:
            _case(f"  tile {rows}: series {i} ({lens[i]} samples)")
            np.testing.assert_array_equal(out[i].cpu().numpy(), exp[i],
                                          err_msg=f"series {i}, tile rows {rows}")
    best = np.unravel_index(np.argmax(exp[big[-1]]), exp[big[-1]].shape)
    print(f"[result] {len(picks)} sampled series bit-exact at tiles 8 and 16; last 2^23 "
          f"series peak S/N {exp[big[-1]][best]:.2f}", flush=True)


# ------------------------------------------------------------------ FFA jerk 2^24


def _jerk_cfg(nsamps: int, bseg_ffa: int):
    from pyloki_amd.config import ParamLimits, PulsarSearchConfig

    lim = ParamLimits.from_taylor((330.0, 336.7), accel=(-30.0, 30.0), jerk=(-2.0, 2.0)).limits
    return PulsarSearchConfig(nsamps=nsamps, tsamp=64e-6, nbins=128, eta=1.0, param_limits=lim,
                              bseg_brute=2048, bseg_ffa=bseg_ffa, ducy_max=0.2, wtsp=1.5,
                              use_fourier=False)


def test_ffa_jerk_2e24_fp32_bf16_vs_oracle(gpu_device):
    """bseg_ffa = 2^21 (N/8): the last level that fits with its input (37.7 GB of FP32
    rows); segment 0 of the last level depends only on the first 2^21 samples, so the
    oracle folds those with the same grids (the grids depend on tseg_brute and the level,
    not on nsamps)."""
    from pyloki_amd import kernels
    from pyloki_amd.ffa import FFAPlan
    from pyloki_amd.simulation.pulse import PulseSignalConfig

    _case("FFA jerk 2^24, bseg_ffa 2^21, FP32 and BF16 storage")
    n, cap = 2**24, 2**21
    cfg = _jerk_cfg(n, cap)
    sub = _jerk_cfg(cap, cap)
    assert cfg.niters_ffa == sub.niters_ffa == 10
    for lvl in range(cfg.niters_ffa + 1):
        assert np.array_equal(cfg.get_param_grid_count(lvl), sub.get_param_grid_count(lvl))
    counts = [int(c) for c in cfg.get_param_grid_count(cfg.niters_ffa)]
    assert counts == [8, 10, 57553]
    ts = PulseSignalConfig(period=3e-3, dt=64e-6, nsamps=n, snr=30.0, ducy=0.08,
                           mod_kwargs={"acc": 10.0, "jerk": 0.5}).generate_analytic(seed=5)
    widths = cfg.score_widths
    plan = FFAPlan(cfg)
    e = torch.from_numpy(ts.ts_e).cuda()
    v = torch.from_numpy(ts.ts_v).cuda()
    fold32 = plan.run(e, v)
    assert fold32.shape[0] == 8
    snr32 = kernels.score_boxcar(fold32[0], widths).cpu().numpy()
    del fold32
    torch.cuda.empty_cache()
    fold16 = plan.run(e, v, dtype=torch.bfloat16)
    snr16 = kernels.score_boxcar(fold16[0], widths).cpu().numpy()
    del fold16
    torch.cuda.empty_cache()
    e0, v0 = ts.ts_e[:cap], ts.ts_v[:cap]
    exp32 = orc.ffa_search_snr(e0, v0, sub).reshape(snr32.shape)
    np.testing.assert_array_equal(snr32, exp32)
    f16, _ = orc.ffa_fold_bf16(e0, v0, sub)
    exp16 = orc.boxcar_snr(orc.fold_norm(orc.bf16_to_f32(f16[0])), widths)
    del f16
    np.testing.assert_array_equal(snr16, exp16)
    snr64 = orc.ffa_search_snr_f64(e0, v0, sub).reshape(snr32.shape)
    peak = float(np.max(np.abs(snr64)))
    big = np.abs(snr64) >= 1.0
    err32 = float(np.max(np.abs(snr32 - snr64))) / peak
    err16 = float(np.max(np.abs(snr16 - snr64))) / peak
    rel32 = float(np.max(np.abs(snr32 - snr64)[big] / np.abs(snr64[big])))
    rel16 = float(np.max(np.abs(snr16 - snr64)[big] / np.abs(snr64[big])))
    print(f"[result] 2^24 jerk, {snr32.shape[0]} trials x {len(widths)} widths, peak S/N "
          f"{peak:.2f}: FP32 max err / peak {err32:.3e} (max rel. err where |S/N| >= 1: "
          f"{rel32:.3e}); BF16 max err / peak {err16:.3e} (max rel. {rel16:.3e})", flush=True)
    assert err32 <= 1e-5
    assert err16 <= 5e-2


# ------------------------------------------------------------------ EP jerk (2^23, 64 segs)


def _host_tree(t) -> po.Tree:
    return po.Tree(t.leaves.cpu().numpy(), t.folds.cpu().numpy(), t.scores.cpu().numpy(),
                   t.backtracks.cpu().numpy())


def _same_tree(g, o, what: str) -> None:
    np.testing.assert_array_equal(g.scores, o.scores, err_msg=what)
    np.testing.assert_array_equal(g.leaves, o.leaves, err_msg=what)
    np.testing.assert_array_equal(g.folds, o.folds, err_msg=what)
    np.testing.assert_array_equal(g.backtracks, o.backtracks, err_msg=what)


@functools.lru_cache(maxsize=1)
def _jerk_ep():
    import bench
    from pyloki_amd.detection.schemes import bound_scheme

    w = bench.EP_JERK
    dyp = bench.ep_jerk_dyp()
    thr = bound_scheme(dyp.nsegments - 1, w["snr_bound"]).astype(np.float64)
    cfg = dyp.cfg
    pcfg = {"poly_order": 3, "nbins": cfg.nbins, "eta": cfg.eta, "branch_max": cfg.branch_max,
            "counts_init": np.array([len(a) for a in dyp.param_arr]),
            "param_limits": cfg.param_limits, "widths": cfg.score_widths,
            "tiling": cfg.tiling_strategy}
    return w, dyp, thr, pcfg


def _side_level(pr, level: int, cap: int):
    """Level ``level`` run by the GPU on the first ``cap`` rows of the search's current tree,
    beside the search (which goes on with its whole tree). Returns the input and the
    output as host trees."""
    from pyloki_amd.prune import DeviceTree

    t = pr.world_tree
    m = min(t.size, cap)
    cut = DeviceTree(t.leaves[:m].contiguous(), t.folds[:m].contiguous(),
                     t.scores[:m].contiguous(), t.backtracks[:m].contiguous())
    seg = pr._segment_rows(pr.scheme.get_segment_idx(level))
    out, st = pr._level(cut, seg, pr._coords(level), float(pr.threshold_scheme[level - 1]))
    return _host_tree(cut), _host_tree(out), st


def test_ep_jerk_2e23_levels(gpu_device):
    from pyloki_amd.prune import Pruning

    _case("EP jerk 2^23 / 64 segments: first 12 levels vs oracle, deep levels from the GPU tree")
    w, dyp, thr, pcfg = _jerk_ep()
    ref, ms = w["ref_seg"], w["max_sugg"]
    nfirst, cap = 12, 8192
    fold = dyp.fold
    exp = po.prune_run(fold, dyp.param_arr, dyp.dparams_actual, dyp.tseg, thr, ref, pcfg, ms,
                       nlevels=nfirst)
    pr = Pruning(dyp, thr, max_sugg=ms, batch_size=1024)
    pr.initialize(ref)
    scheme = po.MiddleOut(dyp.nsegments, ref, dyp.tseg)
    deep_checked = []
    for level in range(1, dyp.nsegments):
        if level > nfirst and level % 8 == 0:
            prev, got, gst = _side_level(pr, level, cap)
            coords = {"init": scheme.get_coord(0), "cur": scheme.get_current_coord(level),
                      "prev": scheme.get_previous_coord(level), "next": scheme.get_coord(level),
                      "add": scheme.get_segment_coord(level)}
            t, st, _ = po.prune_level(prev, fold[scheme.get_segment_idx(level)], coords,
                                      float(thr[level - 1]), pcfg, ms, 1024)
            _same_tree(got, t, f"deep level {level}")
            assert gst["n_leaves"] == st.n_leaves
            deep_checked.append((level, prev.size, t.size))
        pr.execute_iter()
        if level <= nfirst:
            _same_tree(_host_tree(pr.world_tree), exp["trees"][level - 1], f"level {level}")
            assert pr.pstats[level].n_leaves == exp["stats"][level - 1].n_leaves
        if pr.prune_complete:
            break
    print(f"[result] first {nfirst} levels bit-exact; deep levels (level, parents, survivors) "
          f"{deep_checked} bit-exact", flush=True)
    assert len(deep_checked) >= 4


def test_ep_jerk_2e23_level_loop_equals_per_level(gpu_device):
    """The whole 63-level search: the level loop (one library call for every level) gives
    the per-level path's trees and statistics bit for bit; every ascended survivor scores
    at least the last threshold."""
    from pyloki_amd.prune import Pruning

    _case("EP jerk 2^23: level loop vs per-level path, all levels")
    w, dyp, thr, pcfg = _jerk_ep()
    ref, ms = w["ref_seg"], w["max_sugg"]
    pr = Pruning(dyp, thr, max_sugg=ms, batch_size=1024)
    pr.initialize(ref)
    while not pr.prune_complete and pr.prune_level < dyp.nsegments - 1:
        pr.execute_iter()
    pr2 = Pruning(dyp, thr, max_sugg=ms, batch_size=1024)
    pr2.initialize(ref)
    pr2.execute_levels()
    _same_tree(_host_tree(pr2.world_tree), _host_tree(pr.world_tree), "final tree")
    assert [s.n_leaves for s in pr2.pstats] == [s.n_leaves for s in pr.pstats]
    n_leaves = sum(s.n_leaves for s in pr.pstats[1:])
    print(f"[result] {pr.prune_level} levels, {n_leaves} scored leaves, "
          f"{pr.world_tree.size} survivors: loop == per-level", flush=True)
    pr._do_ascend()
    assert np.all(pr.world_tree.scores_ep.cpu().numpy() >= np.float32(thr[pr.prune_level - 1]))


def test_ep_jerk_2e23_finds_pulsar(gpu_device):
    from pyloki_amd.prune import Pruning

    _case("EP jerk 2^23: search result")
    w, dyp, thr, pcfg = _jerk_ep()
    res = Pruning(dyp, thr, max_sugg=w["max_sugg"], batch_size=1024).execute(w["ref_seg"])
    assert len(res["scores"]) > 0
    best = int(np.argmax(res["scores"]))
    f = res["param_sets"][best, -3, 0]
    print(f"[result] {len(res['scores'])} candidates, best S/N {res['scores'][best]:.2f} at "
          f"f = {f:.5f} Hz", flush=True)
    assert abs(f - 500.0) < 0.01


# ------------------------------------------------------------------ EP circular (2^24, 128)


def _circ_pcfg(dyp) -> dict:
    cfg = dyp.cfg
    return {"poly_order": 5, "nbins": cfg.nbins, "eta": cfg.eta, "branch_max": cfg.branch_max,
            "counts_init": np.array([len(a) for a in dyp.param_arr]),
            "param_limits": cfg.param_limits, "widths": cfg.score_widths,
            "tiling": cfg.tiling_strategy, "prop_sig": cfg.propagator_significance,
            "val_sig": cfg.validation_significance, "p_orb_min": cfg.p_orb_min,
            "x_mass_const": cfg.x_mass_const}


def _circ_compare(g, o, what: str) -> None:
    np.testing.assert_array_equal(g.scores, o.scores, err_msg=what)
    np.testing.assert_array_equal(g.backtracks, o.backtracks, err_msg=what)
    np.testing.assert_array_equal(g.folds, o.folds, err_msg=what)
    np.testing.assert_array_equal(g.leaves[:, -1], o.leaves[:, -1], err_msg=what)
    if len(o.leaves):
        scale = np.maximum(np.abs(o.leaves[:, :-1]).max(axis=0, keepdims=True), 1e-300)
        assert np.all(np.abs(g.leaves[:, :-1] - o.leaves[:, :-1]) <= 1e-12 * scale), what


def _complex_compare(g, o, thr: float, what: str) -> int:
    """Complex folds sum in the matrix unit's order, not the oracle's: survivors are paired
    in child order (a row on one side only must score within 1e-5 of the threshold),
    scores and folds agree within 1e-5 of the level's largest."""
    from tests.test_gpu_complex import _align, _close

    pairs, skipped = _align(g, o, thr, what)
    if len(pairs):
        gi, ej = pairs[:, 0], pairs[:, 1]
        _close(g.scores[gi], o.scores[ej], what + " scores")
        _close(g.folds[gi], o.folds[ej], what + " folds")
    return skipped


@pytest.mark.parametrize("use_fourier", [False, True])
def test_ep_circular_2e24_levels(gpu_device, use_fourier):
    """Real folds: bit-exact (leaf values moved along the orbit within 1e-12 of the column
    scale). Complex folds (the reference's default): within the FP32 tolerance."""
    import bench
    from pyloki_amd.prune import Pruning

    _case(f"EP circular 2^24 / 128 segments, use_fourier={use_fourier}: first 20 levels "
          "vs oracle, deep levels from the GPU tree")
    w = bench.EP_CNN
    dyp = bench.ep_circular_dyp(w["test_seed"], use_fourier=use_fourier)
    thr = np.linspace(w["thresholds"][0], w["thresholds"][1], w["nseg"] - 1)
    pcfg = _circ_pcfg(dyp)
    ref, ms = dyp.nsegments // 2, w["max_sugg"]
    nfirst = 20
    fold = dyp.fold
    pr = Pruning(dyp, thr, max_sugg=ms, batch_size=1024)
    pr.initialize(ref)
    scheme = po.MiddleOut(dyp.nsegments, ref, dyp.tseg)
    checked, flags, skipped = [], set(), 0
    for level in range(1, dyp.nsegments):
        side = None
        if level > nfirst and level % 16 == 0 or (use_fourier and level <= nfirst):
            side = _side_level(pr, level, 4096)
        elif level <= nfirst:
            side = (_host_tree(pr.world_tree), None, None)
        pr.execute_iter()
        if side is not None:
            prev, got, gst = side
            if got is None:
                got, gst = _host_tree(pr.world_tree), None
            coords = {"init": scheme.get_coord(0), "cur": scheme.get_current_coord(level),
                      "next": scheme.get_coord(level), "add": scheme.get_segment_coord(level)}
            t, st, _ = po.circ_prune_level(prev, fold[scheme.get_segment_idx(level)], coords,
                                           float(thr[level - 1]), pcfg, ms, 1024)
            if use_fourier:
                skipped += _complex_compare(got, t, float(thr[level - 1]), f"level {level}")
            else:
                _circ_compare(got, t, f"level {level}")
                n_gpu = ((gst["n_leaves"], gst["n_leaves_phy"]) if gst is not None else
                         (pr.pstats[level].n_leaves, pr.pstats[level].n_leaves_phy))
                assert n_gpu == (st.n_leaves, st.n_leaves_phy), level
            flags |= set(np.unique(t.leaves[:, -1, 1]).tolist())
            checked.append((level, prev.size, t.size))
        if pr.prune_complete:
            break
    print(f"[result] use_fourier={use_fourier}: levels (level, parents, survivors) {checked} "
          f"match the oracle from the GPU's previous tree; near-threshold rows on one side "
          f"{skipped}; flags {sorted(flags)}; {pr.prune_level} levels", flush=True)
    assert len(checked) >= nfirst + 4
    if use_fourier:
        return
    # Real folds: the level loop equals the per-level path over all 127 levels.
    pb = Pruning(dyp, thr, max_sugg=ms, batch_size=1024)
    pb.initialize(ref)
    pb.execute_levels()
    _circ_compare(_host_tree(pb.world_tree), _host_tree(pr.world_tree),
                  "level loop vs per-level path")
    assert [(s.n_leaves, s.n_leaves_phy) for s in pb.pstats] == \
        [(s.n_leaves, s.n_leaves_phy) for s in pr.pstats]


def test_ep_circular_2e24_learned_budget_recall(gpu_device):
    """The learned (hybrid) survivor budget against the exhaustive search on held-out
    series of the full configuration, and against the S/N-only budget at the same budget:
    recall of the exhaustive final candidates and scored leaves."""
    import bench
    from pyloki_amd import cnn
    from pyloki_amd.prune import Pruning

    _case("EP circular 2^24: learned budget recall on held-out series")
    w = bench.EP_CNN
    thr = np.linspace(w["thresholds"][0], w["thresholds"][1], w["nseg"] - 1)
    b = w["survey_budget"]
    res = cnn.budget_experiment(bench.ep_circular_dyp, (1, 2, 3, 4), w["test_seed"], thr,
                                w["max_sugg"], (b,), epochs=w["epochs"])
    ranker = res["ranker"]
    tot = {"hybrid": [0, 0, 0], "snr": [0, 0, 0], "cnn": [0, 0, 0]}
    for seed in w["batch_seeds"]:
        dyp = bench.ep_circular_dyp(seed)
        hooks = {"hybrid": cnn.SurvivorBudget(b, ranker, hybrid=True),
                 "snr": cnn.SurvivorBudget(b), "cnn": cnn.SurvivorBudget(b, ranker)}
        r = cnn.compare_searches(lambda: Pruning(dyp, thr, max_sugg=w["max_sugg"],
                                                 batch_size=1024), dyp.nsegments // 2, hooks)
        nf = r["exhaustive"]["n_final"]
        for k in tot:
            tot[k][0] += round(r[k]["recall"] * nf)
            tot[k][1] += nf
            tot[k][2] += r[k]["trials"]
        assert r["hybrid"]["precision"] == 1.0
    rec = {k: v[0] / max(v[1], 1) for k, v in tot.items()}
    print(f"[result] held-out seeds {list(w['batch_seeds'])}, budget {b}: recall "
          f"{ {k: round(v, 4) for k, v in rec.items()} }, scored leaves "
          f"{ {k: v[2] for k, v in tot.items()} }", flush=True)
    assert rec["hybrid"] >= 0.99
    # What the trained network gains over the S/N-only budget at the same budget: at least
    # the same recall with no more scored leaves, and strictly better in one of them.
    assert rec["hybrid"] >= rec["snr"]
    assert tot["hybrid"][2] <= tot["snr"][2]
    assert rec["hybrid"] > rec["snr"] or tot["hybrid"][2] < tot["snr"][2]
