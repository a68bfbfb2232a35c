"""Host side of the EP results (no GPU): the results file (written and read back), the
notebook read-outs built on it, prune_dyp_tree's argument handling, the periodogram's
save / load / plots, and the recapture key of a captured search."""

from __future__ import annotations

from types import SimpleNamespace

import matplotlib
import numpy as np
import pytest

matplotlib.use("Agg")

from pyloki_amd.io.cands import (LEVEL_DTYPE, PruneResultWriter, PruneStats,  # noqa: E402
                                 PruneStatsCollection, read_results)
from pyloki_amd.periodogram import (Periodogram, PruningStatsPlotter,  # noqa: E402
                                    ScatteredPeriodogram)
from pyloki_amd.prune import determine_ref_segs_pareto, prune_dyp_tree  # noqa: E402

RNG = np.random.default_rng(5)


def _stats(nlev: int) -> PruneStatsCollection:
    c = PruneStatsCollection()
    for lv in range(nlev):
        c.append(PruneStats(level=lv, seg_idx=lv + 3, threshold=0.5 * lv, score_min=1.0,
                            score_max=2.0 + lv, n_branches=10 * lv + 1, n_leaves=20 * lv + 1,
                            n_leaves_phy=20 * lv + 1, n_leaves_surv=5 * lv + 1))
    c.timers["branch"] = 0.25
    return c


def _write(path, runs, param_names=("accel", "freq"), nseg=8):
    with PruneResultWriter(path) as w:
        w.write_metadata(list(param_names), nseg, 2**18, np.linspace(1, 4, nseg - 1))
    for name, (ps, sc, se, st) in runs.items():
        with PruneResultWriter(path, mode="a") as w:
            w.write_run_results(name, np.arange(nseg), ps, sc, se, st)


def _run(n, P=2):
    ps = RNG.normal(size=(n, P + 2, 2))
    return ps, RNG.uniform(5, 10, n).astype(np.float32), RNG.uniform(4, 9, n).astype(np.float32)


def test_results_file_round_trip(tmp_path):
    path = tmp_path / "test_pruning_nstages_8_results.npz"
    a, b = _run(12), _run(0)
    _write(path, {"003_00": (*a, _stats(8)), "005_01": (*b, _stats(3))})
    res = read_results(path)
    assert res["attrs"]["param_names"] == ["accel", "freq"]
    assert res["attrs"]["nsegments"] == 8 and res["attrs"]["max_sugg"] == 2**18
    assert list(res["runs"]) == ["003_00", "005_01"]
    r = res["runs"]["003_00"]
    np.testing.assert_array_equal(r["param_sets"], a[0])
    np.testing.assert_array_equal(r["scores"], a[1])
    np.testing.assert_array_equal(r["scores_ep"], a[2])
    assert r["level_stats"].dtype == LEVEL_DTYPE and len(r["level_stats"]) == 8
    back = PruneStatsCollection.from_arrays(r["level_stats"], r["timer_stats"])
    assert back == _stats(8) and back.timers["branch"] == 0.25
    # the archive is a plain npz: numpy reads it without pickles
    with np.load(path, allow_pickle=False) as z:
        assert "runs/005_01/scores" in z.files
    with pytest.raises(ValueError, match="already exists"), \
            PruneResultWriter(path, mode="a") as w:
        w.write_run_results("003_00", np.arange(8), *a, _stats(2))


def test_scattered_periodogram_and_stats_plotter(tmp_path):
    path = tmp_path / "r.npz"
    a, b = _run(30), _run(7)
    _write(path, {"003_00": (*a, _stats(8)), "005_01": (*b, _stats(8))})
    sp = ScatteredPeriodogram.load(path)
    assert sp.n_runs == 2 and len(sp.data) == 37
    assert list(sp.data.columns) == ["accel", "daccel", "freq", "dfreq", "score_ep", "score",
                                     "run_id"]
    np.testing.assert_array_equal(sp.data["freq"].to_numpy()[:30], a[0][:, 1, 0])
    top = sp.get_summary_cands(n=3)
    assert top.startswith("Top candidates:") and top.count("Run:") == 3
    best = sp.get_best_in_each_run()
    assert best.count("Run:") == 2
    assert "No candidates" in sp.get_summary_cands(snr_min=100.0)
    assert len(sp.get_stats()) == 2
    for fig in (sp.plot_correlation("freq", "accel", true_values={"freq": 0.0}),
                sp.plot_scores("scatter"), sp.plot_scores("hist"),
                sp.plot_scores("hist", run_id="003_00")):
        assert fig is not None
    assert "Total runs: 2" in str(sp)
    with pytest.raises(ValueError, match="Invalid score type"):
        sp.get_summary_cands(score_type="x")
    ps = PruningStatsPlotter.load(path)
    assert ps.n_runs == 2 and len(ps.get_level_stats("003_00")) == 8
    assert ps.plot_level_stats() is not None
    with pytest.raises(ValueError, match="9 fields"):
        ps.add_run(np.zeros(3, dtype=[("a", np.int32)]), "x")


def test_load_rejects_other_files(tmp_path):
    path = tmp_path / "x.npz"
    np.savez(path, a=np.arange(3))
    with pytest.raises(ValueError, match="Not a valid pruning results file"):
        ScatteredPeriodogram.load(path)


def _dyp_stub(nseg=16):
    return SimpleNamespace(nsegments=nseg, cfg=SimpleNamespace(param_names=["accel", "freq"]))


def test_prune_dyp_tree_arguments(tmp_path):
    """prune.py:836-856: n_workers type and range, n_runs range and precedence over
    ref_segs, one of them required, existing output files refused; all before any search."""
    d = _dyp_stub()
    thr = np.ones(15)
    with pytest.raises(TypeError, match="n_workers must be an integer"):
        prune_dyp_tree(d, thr, n_runs=2, outdir=tmp_path, n_workers=2.0)
    with pytest.raises(ValueError, match="n_workers must be greater than 0"):
        prune_dyp_tree(d, thr, n_runs=2, outdir=tmp_path, n_workers=0)
    with pytest.raises(ValueError, match="n_runs must be between 1 and 16"):
        prune_dyp_tree(d, thr, n_runs=17, ref_segs=[3], outdir=tmp_path / "a")
    with pytest.raises(ValueError, match="Either n_runs or ref_segs"):
        prune_dyp_tree(d, thr, outdir=tmp_path / "b")
    (tmp_path / "c").mkdir()
    (tmp_path / "c" / "test_pruning_nstages_16_results.npz").write_bytes(b"")
    with pytest.raises(FileExistsError):
        prune_dyp_tree(d, thr, n_runs=2, outdir=tmp_path / "c")
    assert determine_ref_segs_pareto(16, 1) == [8]
    segs = determine_ref_segs_pareto(16, 4)
    assert len(segs) == 4 and segs == sorted(segs)


def test_periodogram_save_load_and_plots(tmp_path):
    acc = np.linspace(-1, 1, 5)
    freq = np.linspace(199, 201, 40)
    width = np.array([1, 2, 4])
    snrs = RNG.normal(size=(5, 40, 3))
    snrs[2, 17, 1] = 30.0
    pg = Periodogram(params={"width": width, "accel": acc, "freq": freq}, snrs=snrs, tobs=10.0)
    assert pg.param_names == ["accel", "freq", "width"]
    best = pg.find_best_params()
    assert best["snr"] == 30.0 and best["freq"] == freq[17] and best["width"] == 2
    assert pg.get_indices_summary({"freq": freq[17], "accel": acc[2]}).startswith(
        "True param indices: (2, 17)")
    pg.save(tmp_path / "p.npz")
    back = Periodogram.load(tmp_path / "p.npz")
    np.testing.assert_array_equal(back.snrs, snrs)
    assert back.tobs == 10.0 and back.param_names == pg.param_names
    sl = pg.get_slice({"width": 1})
    assert sl.dims == ("accel", "freq") and sl.values[2, 17] == 30.0
    assert pg.plot_1d("freq") is not None
    assert pg.plot_2d("freq", "accel", fixed_param_idx={"width": 1}) is not None
    with pytest.raises(ValueError, match="Invalid parameter"):
        pg.plot_1d("jerk")
    with pytest.raises(ValueError, match="Missing required"):
        Periodogram(params={"freq": freq}, snrs=snrs[0, :, 0], tobs=1.0)


def test_captured_search_recaptures_when_the_budget_changes():
    """The capture key covers the loop buffers and a survivor budget's scratch pointers,
    budget and level flags (all baked into the graph at capture)."""
    from pyloki_amd.cnn import SurvivorBudget
    from pyloki_amd.graph import CapturedSearch
    from pyloki_amd.prune import EpBudget

    hook = SurvivorBudget(100)
    hook._s = EpBudget()
    hook._s.budget = 100
    hook._s.keys = 4096
    buf = SimpleNamespace(s=SimpleNamespace(hist=1, cap_child=2, c_score=3))
    cs = object.__new__(CapturedSearch)
    cs.pr = SimpleNamespace(_loop=buf, level_hook=hook)
    k0 = cs._buffer_ptrs()
    hook._s.budget = 50
    k1 = cs._buffer_ptrs()
    hook._s.keys = 8192
    k2 = cs._buffer_ptrs()
    hook.level_mask = np.ones(4, dtype=np.uint8)
    k3 = cs._buffer_ptrs()
    assert len({k0, k1, k2, k3}) == 4
