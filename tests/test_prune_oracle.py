"""Pin the EP (Taylor pruning) oracle: branching tiles, resolve, survivor sets, overflow.

CPU only; sizes chosen so the whole file runs in seconds.
"""

from __future__ import annotations

import math
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle import oracle as orc
from oracle import prune_oracle as po
from pyloki_amd.config import ParamLimits, PulsarSearchConfig
from pyloki_amd.simulation.pulse import PulseSignalConfig

RNG = np.random.default_rng(77)


def _leaves(n: int, P: int = 2) -> np.ndarray:
    lv = np.zeros((n, P + 2, 2))
    lv[:, 0, 0] = RNG.uniform(-30, 30, n)          # accel (or jerk)
    lv[:, 0, 1] = RNG.uniform(0.5, 20, n)
    if P == 3:
        lv[:, 1, 0] = RNG.uniform(-30, 30, n)
        lv[:, 1, 1] = RNG.uniform(0.5, 20, n)
    lv[:, -3, 0] = RNG.uniform(-50, 50, n)         # velocity
    lv[:, -3, 1] = RNG.uniform(100, 4000, n)
    lv[:, -1, 0] = RNG.uniform(195, 205, n)        # f0
    return lv


def test_middle_out_scheme():
    s = po.MiddleOut(8, 3, 2.0)
    assert s.data.tolist() == [3, 2, 4, 1, 5, 0, 6, 7]
    assert s.get_coord(0) == (7.0, 1.0)
    assert s.get_coord(2) == (7.0, 3.0)          # segments 2..4 -> centre 3.5*2, half 1.5*2
    assert s.get_segment_coord(1) == (5.0, 1.0)
    assert s.get_current_coord(2) == (s.get_coord(1)[0], s.get_coord(2)[1])
    with pytest.raises(ValueError, match="ref_idx"):
        po.MiddleOut(8, 8, 1.0)


@pytest.mark.parametrize("P", [2, 3])
def test_branch_tiles_parent_exactly(P):
    lv = _leaves(300, P)
    children, origins = po.branch_batch(lv, (10.0, 40.0), 64, 1.0, P, 64)
    assert np.all(np.diff(origins) >= 0)  # parent-major
    for i in range(len(lv)):
        ch = children[origins == i]
        for j in range(P):
            vals = np.unique(ch[:, j, 0])
            da = ch[0, j, 1]
            n = len(vals)
            pc, dc = lv[i, j, 0], lv[i, j, 1]
            if n == 1 and da == dc:
                assert vals[0] == pc
                continue
            assert n == max(1, math.ceil(dc / da - 1e-9)) or n == round(dc / da)
            # children tile [pc - dc/2, pc + dc/2] exactly: width n * da == dc
            assert n * da == pytest.approx(dc, rel=1e-12)
            assert vals[0] - da / 2 == pytest.approx(pc - dc / 2, rel=1e-12, abs=1e-12)
            assert vals[-1] + da / 2 == pytest.approx(pc + dc / 2, rel=1e-12, abs=1e-12)
        # odometer order: last branched parameter fastest
        if len(ch) > 1:
            key = [tuple(row) for row in ch[:, :P, 0]]
            assert key == sorted(key, key=lambda t: tuple(t[k] for k in range(P)))


def test_branch_max_error():
    lv = _leaves(4)
    lv[:, 0, 1] = 1e6
    with pytest.raises(ValueError, match="increase branch_max"):
        po.branch_batch(lv, (10.0, 200.0), 64, 1.0, 2, 16)


def test_resolve_vs_exact_rational():
    P = 2
    lv = _leaves(400, P)
    limits = np.array([[-50.0, 50.0], [190.0, 210.0]])
    counts = np.array([17, 4001])
    coords = {"add": (30.0, 1.0), "cur": (20.0, 5.0), "init": (16.0, 1.0)}
    idx, phase = po.resolve_batch(lv, coords["add"], coords["cur"], coords["init"], counts,
                                  limits, 64)
    C = Fr(po.C_VAL)
    checked = 0
    for n in range(len(lv)):
        a, v, d0, f0 = (Fr(lv[n, 0, 0]), Fr(lv[n, 1, 0]), Fr(lv[n, 2, 0]), Fr(lv[n, 3, 0]))

        def at(dt):
            dt = Fr(dt)
            return a, v + a * dt, d0 + v * dt + a * dt * dt / 2

        a_add, v_add, d_add = at(30.0 - 20.0)
        _, v_init, d_init = at(16.0 - 20.0)
        freq = f0 * (1 - (v_add - v_init) / C)
        delay = (d_add - d_init) / C
        x = (Fr(30.0 - 16.0) - delay) * f0
        pos = (x - math.floor(x)) * 64
        ra = (a_add - Fr(-50)) * Fr(17) / Fr(100) + Fr(1, 10**6)
        rf = (freq - Fr(190)) * Fr(4001) / Fr(20) + Fr(1, 10**6)
        dist = min(abs(ra - round(ra)), abs(rf - round(rf)),
                   abs(pos + Fr(1, 2) - round(pos + Fr(1, 2))))
        if dist < 1e-6:
            continue
        assert idx[n, 0] == min(max(math.floor(ra), 0), 16)
        assert idx[n, 1] == min(max(math.floor(rf), 0), 4000)
        assert orc.shift_int(phase[n], 64) == math.floor(pos + Fr(1, 2)) % 64
        checked += 1
    assert checked > 390


def _small_search(nsamps=2**17, nseg_ffa=8, seed=4):
    tsamp = 64e-6
    lim = ParamLimits.from_taylor((195.0, 205.0), accel=(-50.0, 50.0)).limits
    cfg = PulsarSearchConfig(nsamps=nsamps, tsamp=tsamp, nbins=32, eta=1.0, param_limits=lim,
                             bseg_brute=nsamps // 256, bseg_ffa=nsamps // nseg_ffa,
                             prune_poly_order=2, ducy_max=0.3, wtsp=1.5, use_fourier=False,
                             branch_max=32)
    sim = PulseSignalConfig(period=1 / 200.37, dt=tsamp, nsamps=nsamps, snr=30.0, ducy=0.1,
                            mod_kwargs={"acc": 12.0})
    ts = sim.generate_analytic(seed=seed)
    fold, pa = orc.ffa_fold(ts.ts_e, ts.ts_v, cfg)
    counts = [len(a) for a in pa]
    dyp_fold = fold.reshape(fold.shape[0], *counts, 2, cfg.nbins)
    pcfg = {"poly_order": 2, "nbins": cfg.nbins, "eta": cfg.eta, "branch_max": cfg.branch_max,
            "counts_init": np.array(counts), "param_limits": cfg.param_limits,
            "widths": cfg.score_widths, "tiling": "aggressive"}
    return cfg, dyp_fold, pa, pcfg


def test_level_survivors_equal_bruteforce_without_overflow():
    cfg, dyp_fold, pa, pcfg = _small_search()
    tseg = cfg.tseg_ffa
    res = po.prune_run(dyp_fold, pa, cfg.get_dparams_actual(cfg.niters_ffa), tseg,
                       np.full(7, 3.0), ref_seg=3, cfg=pcfg, max_sugg=2**20, nlevels=2)
    scheme = res["scheme"]
    # Level 1 from the seed: brute force every child independently.
    leaves = po.taylor_seed(pa, cfg.get_dparams_actual(cfg.niters_ffa), 2)
    folds0 = dyp_fold[3].reshape(-1, 2, cfg.nbins)
    coords = {"init": scheme.get_coord(0), "cur": scheme.get_current_coord(1),
              "next": scheme.get_coord(1), "add": scheme.get_segment_coord(1)}
    ch, org = po.branch_batch(leaves, coords["cur"], 32, 1.0, 2, 32)
    idx, ph = po.resolve_batch(ch, coords["add"], coords["cur"], coords["init"],
                               pcfg["counts_init"], cfg.param_limits, 32)
    seg = dyp_fold[scheme.get_segment_idx(1)]
    sc = np.empty(len(ch), dtype=np.float32)
    for n in range(len(ch)):
        row = seg[idx[n, 0], idx[n, 1]]
        s = orc.shift_int(ph[n], 32)
        comb = (folds0[org[n]] + np.roll(row, s, axis=-1)).astype(np.float32)
        sc[n] = orc.snr_score_batch(comb[None], pcfg["widths"])[0]
    keep = np.where(sc.astype(np.float64) >= 3.0)[0]
    np.testing.assert_array_equal(res["survivors"][0], keep)
    np.testing.assert_array_equal(res["trees"][0].scores, sc[keep])
    assert res["stats"][0].n_leaves == len(ch)


def _sorted_overflow(valid_scores, batch_scores, size, cur):
    """Sort-based restatement of prune_on_overload + add_batch (world_tree.py:482-551)."""
    allv = np.concatenate([valid_scores, batch_scores]).astype(np.float32)
    total = len(allv)
    srt = np.sort(allv)
    kth = srt[total - size]
    mid = srt[total - (total // 2) - 1]
    eff = max(cur, math.nextafter(float(kth), math.inf), math.nextafter(float(mid), math.inf))
    kept = [s for s in valid_scores if float(s) >= eff]
    space = size - len(kept)
    add = [s for s in batch_scores if float(s) >= eff][:space]
    return np.array(kept + add, dtype=np.float32), eff


def test_overflow_rule_matches_sort_restatement():
    for trial in range(50):
        size = int(RNG.integers(5, 40))
        buf = po.TreeBuffer(size, 2, 8)
        valid = RNG.normal(5, 1, size).astype(np.float32)
        if trial % 3 == 0:
            valid[: size // 3] = valid[size // 3]  # ties
        z = np.zeros((size, 4, 2))
        buf.add_batch(z, np.zeros((size, 2, 8), np.float32), valid,
                      np.zeros((size, 4), np.int32), np.arange(size), 0.0)
        nb = int(RNG.integers(1, 60))
        batch = RNG.normal(5.2, 1, nb).astype(np.float32)
        cur = float(RNG.uniform(3, 5))
        eff = buf.add_batch(np.zeros((nb, 4, 2)), np.zeros((nb, 2, 8), np.float32), batch,
                            np.zeros((nb, 4), np.int32), np.arange(nb), cur)
        exp_scores, exp_eff = _sorted_overflow(valid, batch, size, cur)
        assert eff == exp_eff
        np.testing.assert_array_equal(buf.scores[: buf.valid], exp_scores)


def test_search_recovers_injected_signal():
    cfg, dyp_fold, pa, pcfg = _small_search()
    res = po.prune_run(dyp_fold, pa, cfg.get_dparams_actual(cfg.niters_ffa), cfg.tseg_ffa,
                       np.array([2.0, 2.5, 3.0, 3.5, 4.0, 4.5, 5.0]), ref_seg=3, cfg=pcfg,
                       max_sugg=2**16)
    tree = res["tree"]
    assert tree.size > 0
    best = int(np.argmax(tree.scores))
    rep = res["report"][best]
    assert abs(rep[-3, 0] - 200.37) < 0.02          # frequency
    # the injected acceleration lies inside the best leaf's acceleration cell
    assert abs(tree.leaves[best, 0, 0] - 12.0) <= tree.leaves[best, 0, 1] / 2
    assert tree.scores[best] > 15


def test_fixed_grid_coords_by_hand_and_product_scheme():
    """snail.py:286-345 fixed coordinates: t_init with the half-width from t_init to the far
    edge of the segments so far, less the initial half-width; the report keeps the last
    level's t0. The product's MiddleOutScheme gives the same coordinates."""
    from pyloki_amd.utils.snail import MiddleOutScheme

    s = po.MiddleOut(8, 3, 1.0)  # order 3, 2, 4, 1, 5, 0, 6, 7
    assert s.get_current_coord(0, False) == (3.5, 0.5)
    # level 1: segments 2..3 -> edges 2 and 4 -> max(1.5, 0.5) - 0.5
    assert s.get_current_coord(1, False) == (3.5, 1.0)
    assert s.get_current_coord(2, False) == (3.5, 1.0)   # 2..4: edges 2, 5 -> 1.5
    assert s.get_current_coord(7, False) == (3.5, 4.0)   # 0..7: max(3.5, 4.5) - 0.5
    assert s.get_previous_coord(1, False) == s.get_coord(0)
    assert s.get_previous_coord(3, False) == s.get_current_coord(2, False)
    assert s.get_report_coord(False) == (4.0, 4.0)
    assert s.get_report_coord(True) == s.get_coord(7)
    for nseg, ref, ts in ((8, 3, 1.0), (11, 0, 0.7), (11, 10, 2.5), (16, 7, 1.3)):
        o, p = po.MiddleOut(nseg, ref, ts), MiddleOutScheme(nseg, ref, ts)
        for mv in (True, False):
            for lv in range(nseg):
                assert p.get_current_coord(lv, mv) == o.get_current_coord(lv, mv)
            for lv in range(1, nseg + 1):
                assert p.get_previous_coord(lv, mv) == o.get_previous_coord(lv, mv)
            assert p.get_report_coord(mv) == o.get_report_coord(mv)


def test_fixed_resolve_vs_exact_rational():
    """taylor.py:244-277: velocity and delay at t_add are the whole polynomial values (the
    moving grid's init offsets are not subtracted), phase from t_add - t_init."""
    P = 2
    lv = _leaves(400, P)
    lv[:, 2, 0] = 0.0  # the fixed grid keeps d0 = 0 at t_init
    limits = np.array([[-50.0, 50.0], [190.0, 210.0]])
    counts = np.array([17, 4001])
    idx, phase = po.fixed_resolve_batch(lv, (30.0, 1.0), (16.0, 1.0), counts, limits, 64)
    C = Fr(po.C_VAL)
    checked = 0
    for n in range(len(lv)):
        a, v, f0 = Fr(lv[n, 0, 0]), Fr(lv[n, 1, 0]), Fr(lv[n, 3, 0])
        dt = Fr(30.0 - 16.0)
        v_add, d_add = v + a * dt, v * dt + a * dt * dt / 2
        freq = f0 * (1 - v_add / C)
        x = (dt - d_add / C) * f0
        pos = (x - math.floor(x)) * 64
        ra = (a - Fr(-50)) * Fr(17) / Fr(100) + Fr(1, 10**6)
        rf = (freq - Fr(190)) * Fr(4001) / Fr(20) + Fr(1, 10**6)
        dist = min(abs(ra - round(ra)), abs(rf - round(rf)),
                   abs(pos + Fr(1, 2) - round(pos + Fr(1, 2))))
        if dist < 1e-6:
            continue
        assert idx[n, 0] == min(max(math.floor(ra), 0), 16)
        assert idx[n, 1] == min(max(math.floor(rf), 0), 4000)
        assert orc.shift_int(phase[n], 64) == math.floor(pos + Fr(1, 2)) % 64
        checked += 1
    assert checked > 390
    # the moving resolve from the same reference time subtracts the init velocity: with
    # velocities of a few hundred km/s the frequency bins differ
    big = lv.copy()
    big[:, 1, 0] = RNG.uniform(1e5, 3e5, len(big))
    im, _ = po.resolve_batch(big, (30.0, 1.0), (16.0, 1.0), (16.0, 1.0), counts, limits, 64)
    fx, _ = po.fixed_resolve_batch(big, (30.0, 1.0), (16.0, 1.0), counts, limits, 64)
    assert np.mean(im[:, 1] != fx[:, 1]) > 0.9


def test_fixed_resolve_circular_and_chebyshev_identities():
    """circular.py:531-590 and chebyshev.py:261-300: the fixed resolve is the moving one
    from t_cur = t_init without the init velocity and delay, and the circular fixed resolve
    writes no basis flags."""
    from oracle import cheby_oracle as co

    n = 300
    lv = np.zeros((n, 7, 2))
    lv[:, 0, 0] = RNG.uniform(-1e-6, 1e-6, n)
    lv[:, 1, 0] = RNG.uniform(-3e-3, 3e-3, n)
    lv[:, 2, 0] = RNG.uniform(-0.3, 0.3, n)
    lv[:, 3, 0] = RNG.uniform(-30, 30, n)
    lv[:, 4, 0] = RNG.uniform(1e5, 3e5, n)  # velocities large enough to move a bin
    lv[:, :5, 1] = np.abs(lv[:, :5, 0]) * RNG.uniform(0.01, 2, (n, 5))
    lv[:, 6, 0] = RNG.uniform(199, 201, n)
    lv[:, 6, 1] = 7.0
    limits = np.array([[-1, 1], [-1, 1], [-1, 1], [-40, 40], [190, 210.0]])
    counts = np.array([1, 1, 1, 9, 2001])
    init, add = (16.0, 1.0), (30.0, 1.0)
    before = lv.copy()
    fi, fp = po.circ_fixed_resolve_batch(lv, add, init, counts, limits, 64, 2.0)
    np.testing.assert_array_equal(lv, before)  # no flag writes
    acc, freq_mov, _, masks = po.circ_resolve_values(lv, add, init, init, 2.0)
    assert masks[0].any() and masks[2].any()
    d_add = po.circ_shift_all(lv[:, :-1, 0], 14.0, *masks)
    f0 = lv[:, -1, 0]
    f_fixed = f0 * (1 - d_add[:, -2] / po.C_VAL)
    np.testing.assert_array_equal(fi, po.nearest_2d(acc, f_fixed, counts, limits))
    np.testing.assert_array_equal(fp, po.phase_idx_vec(14.0, f0, 64, d_add[:, -1] / po.C_VAL))
    moving = po.nearest_2d(acc, freq_mov, counts, limits)
    assert np.mean(moving[:, -1] != fi[:, -1]) > 0.9
    # Chebyshev: the fixed resolve's implied velocity exceeds the moving one's by v(t_init)
    P = 2
    cl = np.zeros((n, P + 2, 2))
    cl[:, 0, 0] = RNG.uniform(-3, 3, n)
    cl[:, 1, 0] = RNG.uniform(-50, 50, n)
    cl[:, 2, 0] = RNG.uniform(-2, 2, n)
    cl[:, 3, 0] = RNG.uniform(195, 205, n)
    lim2 = np.array([[-50.0, 50.0], [190.0, 210.0]])
    cnt2 = np.array([17, 4001])
    cur = (16.0, 9.0)
    fi2, fp2 = co.cheby_fixed_resolve_batch(cl, add, cur, init, cnt2, lim2, 64)
    d_add = co.cheby_param_shift(cl[:, :-1, 0], 16.0, 9.0, 30.0)
    d_in = co.cheby_param_shift(cl[:, :-1, 0], 16.0, 9.0, 16.0)
    f_fix = cl[:, -1, 0] * (1 - d_add[:, -2] / po.C_VAL)
    np.testing.assert_array_equal(fi2, po.nearest_2d(d_add[:, -3], f_fix, cnt2, lim2))
    im2, pm2 = co.cheby_resolve_batch(cl, add, cur, init, cnt2, lim2, 64)
    d_mov = (d_add[:, -1] - d_in[:, -1]) / po.C_VAL
    np.testing.assert_array_equal(pm2, po.phase_idx_vec(14.0, cl[:, -1, 0], 64, d_mov))
    np.testing.assert_array_equal(fp2, po.phase_idx_vec(14.0, cl[:, -1, 0], 64,
                                                        d_add[:, -1] / po.C_VAL))


def test_fixed_grid_search_recovers_injected_signal():
    """The whole fixed-grid run (moving=False) finds the injected pulsar as the moving one
    does; its survivors stay referenced to the reference segment."""
    cfg, dyp_fold, pa, pcfg = _small_search()
    thr = np.array([2.0, 2.5, 3.0, 3.5, 4.0, 4.5, 5.0])
    dp = cfg.get_dparams_actual(cfg.niters_ffa)
    res = po.prune_run(dyp_fold, pa, dp, cfg.tseg_ffa, thr, ref_seg=3, cfg=pcfg, max_sugg=2**16,
                       moving=False)
    mov = po.prune_run(dyp_fold, pa, dp, cfg.tseg_ffa, thr, ref_seg=3, cfg=pcfg, max_sugg=2**16)
    tree = res["tree"]
    assert tree.size > 0
    best = int(np.argmax(tree.scores))
    assert abs(res["report"][best][-3, 0] - 200.37) < 0.02
    assert abs(tree.leaves[best, 0, 0] - 12.0) <= tree.leaves[best, 0, 1] / 2
    assert tree.scores[best] > 15
    assert np.all(tree.leaves[:, -2, 0] == 0.0)   # d0 stays 0: never re-centred
    assert float(tree.scores.max()) >= 0.85 * float(mov["tree"].scores.max())
