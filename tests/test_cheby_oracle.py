"""Pin the Chebyshev-basis EP oracle.

The reference's own known answers (``tests/test_maths.py:55-149``), the domain-transfer
identity of ``tests/cheby.py`` and the generalised-polynomial matrix of
``tests/mat_inv_tests.py`` pin the tables and transforms. numpy's Chebyshev series pin the
parameter shift (derivatives at a time). A level is checked against a child-by-child brute
force, the branching pattern against the counts of an unpruned search, and the search's
candidate against the Taylor-basis search of the same data. CPU only, seconds.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
from numpy.polynomial import Chebyshev, Polynomial
from numpy.polynomial import chebyshev as npc

from oracle import cheby_oracle as co
from oracle import oracle as orc
from oracle import prune_oracle as po
from pyloki_amd.config import ParamLimits, PulsarSearchConfig
from pyloki_amd.simulation.pulse import PulseSignalConfig

RNG = np.random.default_rng(123)


def test_connection_coefficients_known_answers():
    # test_maths.py:59-91
    assert co.conn_s(0, 0) == 1.0
    assert co.conn_s(2, 0) == 0.5 and co.conn_s(2, 2) == 0.5
    assert co.conn_s(3, 1) == 0.75 and co.conn_s(3, 3) == 0.25
    assert co.conn_r(2, 0) == -1.0 and co.conn_r(2, 2) == 2.0
    # S and R are the power <-> Chebyshev conversions: R = S^-1, both match numpy's
    for K in (2, 4, 6):
        S, R = co.matrix_s(K), co.matrix_r(K)
        np.testing.assert_allclose(S @ R, np.eye(K + 1), atol=1e-13)
        for k in range(K + 1):
            e = np.zeros(K + 1)
            e[k] = 1.0
            np.testing.assert_allclose(S[k], np.pad(npc.poly2cheb(e), (0, K + 1))[:K + 1],
                                       atol=1e-14)
            np.testing.assert_allclose(R[k], np.pad(npc.cheb2poly(e), (0, K + 1))[:K + 1],
                                       atol=1e-14)


def test_taylor_to_cheby_closed_form():
    # test_maths.py:93-109
    d = np.array([0.5, 2.3, 1500.0, 1e6, 1e4])
    ts = 4.2
    f = math.factorial
    exp = np.array([
        d[0] * ts**4 / (8 * f(4)),
        d[1] * ts**3 / (4 * f(3)),
        0.5 * ((d[2] * ts**2 / f(2)) + (d[0] * ts**4 / f(4))),
        d[3] * ts + (0.75 * d[1] * ts**3 / f(3)),
        d[4] + d[2] * ts**2 / (2 * f(2)) + 3 * d[0] * ts**4 / (8 * f(4)),
    ])
    np.testing.assert_allclose(co.taylor_to_cheby(d[None], ts)[0], exp, rtol=1e-14)


def test_cheby_to_taylor_closed_form():
    # test_maths.py:111-121
    a = np.array([0.5, 2.3, 1500.0, 1e6, 1e4])
    ts = 4.2
    exp = np.array([192 * a[0] / ts**4, 24 * a[1] / ts**3, 4 * (a[2] - 4 * a[0]) / ts**2,
                    (a[3] - 3 * a[1]) / ts, a[4] - a[2] + a[0]])
    np.testing.assert_allclose(co.cheby_to_taylor(a[None], ts)[0], exp, rtol=1e-14)


@pytest.mark.parametrize("kmax", [2, 4, 6])
def test_roundtrip_and_evaluation(kmax):
    # test_maths.py:123-149
    d = RNG.random((5, kmax + 1))
    ts, tc = 1.5, 4.2
    a = co.taylor_to_cheby(d, ts)
    np.testing.assert_allclose(co.cheby_to_taylor(a, ts), d, rtol=1e-12, atol=1e-14)
    t = np.linspace(tc - ts, tc + ts, 11)
    for r in range(5):
        taylor = Polynomial(d[r, ::-1] / [math.factorial(k) for k in range(kmax + 1)])
        cheb = Chebyshev(a[r, ::-1], domain=[tc - ts, tc + ts])
        np.testing.assert_allclose(cheb(t), taylor(t - tc), rtol=1e-12)
    full = np.stack([d, RNG.random((5, kmax + 1))], axis=-1)
    back = co.cheby_to_taylor_full(co.taylor_to_cheby_full(full, ts), ts)
    np.testing.assert_allclose(back[..., 0], full[..., 0], rtol=1e-12, atol=1e-14)
    # widths: each Taylor width alone maps to |S| scaled, and back through |R|
    np.testing.assert_allclose(co.taylor_to_cheby_errors(full[..., 1], ts),
                               co.taylor_to_cheby_full(full, ts)[..., 1], rtol=1e-15)


@pytest.mark.parametrize("p,q", [(1.0, 0.0), (2.0, 0.0), (1.0, 0.37), (0.6, -1.3), (3.5, 2.25)])
def test_transform_coefficient_identity(p, q):
    # tests/cheby.py:66-82: T_n(p x + q) = sum_k C_{n,k}(p, q) T_k(x)
    x = np.linspace(-1, 1, 9)
    for n in range(7):
        e = np.zeros(n + 1)
        e[n] = 1.0
        lhs = npc.chebval(p * x + q, e)
        c = np.array([co.trans_c(n, k, p, q) for k in range(n + 1)])
        rhs = npc.chebval(x, c)
        np.testing.assert_allclose(rhs, lhs, rtol=1e-12, atol=1e-12 * max(1, abs(lhs).max()))
        assert all(co.trans_c(n, k, p, q) == 0.0 for k in range(n + 1, 8))


def test_transform_matrix_matches_generalised_polynomials():
    # tests/mat_inv_tests.py:93-117: C = M1 inv(M2), M rows = T_n((t - tc)/ts) in powers of t
    K, (tc1, ts1), (tc2, ts2) = 4, (5.0, 1.0), (10.0, 2.0)

    def gen(tc, ts):
        rows = []
        for n in range(K + 1):
            coef = Chebyshev.basis(n, domain=[tc - ts, tc + ts]).convert(kind=Polynomial).coef
            rows.append(np.pad(coef, (0, K + 1 - len(coef))))
        return np.array(rows)

    ref = gen(tc1, ts1) @ np.linalg.inv(gen(tc2, ts2))
    asc = co.transform_matrix(K, tc1, ts1, tc2, ts2, descending=False)
    np.testing.assert_allclose(asc, ref, rtol=1e-9, atol=1e-9)
    desc = co.transform_matrix(K, tc1, ts1, tc2, ts2)
    np.testing.assert_array_equal(desc, asc[::-1, ::-1])


@pytest.mark.parametrize("tiling", ["aggressive", "conservative", "quadrature"])
def test_shift_cheby_full_keeps_the_polynomial(tiling):
    K = 3
    a = np.stack([RNG.normal(size=(6, K + 1)), RNG.uniform(0.1, 1, (6, K + 1))], axis=-1)
    cur, nxt = (3.0, 1.5), (3.75, 2.25)
    out = co.shift_cheby_full(a, nxt, cur, tiling)
    t = np.linspace(2.0, 5.0, 13)
    for r in range(6):
        f_old = Chebyshev(a[r, ::-1, 0], domain=[cur[0] - cur[1], cur[0] + cur[1]])
        f_new = Chebyshev(out[r, ::-1, 0], domain=[nxt[0] - nxt[1], nxt[0] + nxt[1]])
        np.testing.assert_allclose(f_new(t), f_old(t), rtol=1e-12, atol=1e-12)
    C = co.transform_matrix(K, cur[0], cur[1], nxt[0], nxt[1])
    if tiling == "aggressive":
        exp = a[..., 1] * np.abs(np.diag(C))
    elif tiling == "conservative":
        exp = a[..., 1] @ np.abs(C)
    else:
        exp = np.sqrt(a[..., 1] ** 2 @ C**2)
    np.testing.assert_allclose(out[..., 1], exp, rtol=1e-13)


def test_param_shift_is_the_series_derivatives():
    K = 4
    a = RNG.normal(size=(5, K + 1)) * np.array([1e-3, 1e-2, 1.0, 10.0, 100.0])
    t0, ts = 2.5, 1.25
    for t_eval in (0.3, 2.5, 3.9, 6.0):
        d = co.cheby_param_shift(a, t0, ts, t_eval)
        for r in range(5):
            s = Chebyshev(a[r, ::-1], domain=[t0 - ts, t0 + ts])
            exp = np.array([s.deriv(k)(t_eval) if k else s(t_eval) for k in range(K, -1, -1)])
            np.testing.assert_allclose(d[r], exp, rtol=1e-9, atol=1e-9 * np.abs(exp).max())


def test_seed_closed_form():
    pa = [np.array([-10.0, 0.0, 10.0]), np.array([199.0, 201.0])]
    dp = np.array([5.0, 0.01])
    ts = 0.8
    lv = co.cheby_seed(pa, dp, 2, (3.0, ts))
    assert lv.shape == (6, 4, 2)
    acc = np.repeat(pa[0], 2)
    f0 = np.tile(pa[1], 3)
    dv = 0.01 * (co.C_VAL / f0)
    np.testing.assert_allclose(lv[:, 0, 0], acc * ts**2 / 4, rtol=1e-15)
    np.testing.assert_allclose(lv[:, 1, 0], 0.0)
    np.testing.assert_allclose(lv[:, 2, 0], acc * ts**2 / 4, rtol=1e-15)
    np.testing.assert_allclose(lv[:, 0, 1], 5.0 * ts**2 / 4, rtol=1e-15)
    np.testing.assert_allclose(lv[:, 1, 1], dv * ts, rtol=1e-15)
    np.testing.assert_allclose(lv[:, 2, 1], 5.0 * ts**2 / 4, rtol=1e-15)
    np.testing.assert_array_equal(lv[:, 3, 0], f0)
    np.testing.assert_array_equal(lv[:, 3, 1], 0.0)


@pytest.mark.parametrize("P", [2, 3])
def test_branch_tiles_the_transformed_parent(P):
    n = 40
    lv = np.zeros((n, P + 2, 2))
    lv[:, :P + 1, 0] = RNG.normal(size=(n, P + 1)) * 1e3
    lv[:, :P, 1] = RNG.uniform(50.0, 3000.0, (n, P))
    lv[:, P, 1] = RNG.uniform(0, 1, n)
    lv[:, -1, 0] = RNG.uniform(195, 205, n)
    prev, cur = (4.0, 2.0), (4.0, 3.0)
    ch, org = co.cheby_branch_batch(lv, cur, prev, 64, 1.0, P, 256, "aggressive")
    tr = co.shift_cheby_full(lv[:, :-1], cur, prev, "aggressive")
    dnew = (1.0 / 64) * co.C_VAL / lv[:, -1, 0]
    assert np.all(np.diff(org) >= 0)
    for i in range(n):
        kids = ch[org == i]
        for j in range(P):
            shift = abs(tr[i, j, 1] - dnew[i]) * (64 * (lv[i, -1, 0] / co.C_VAL))
            vals = np.unique(kids[:, j, 0])
            if shift < 1.0 - 1e-6:
                np.testing.assert_array_equal(vals, [tr[i, j, 0]])
                assert np.all(kids[:, j, 1] == tr[i, j, 1])
                continue
            m = max(1, math.ceil(tr[i, j, 1] / dnew[i] - 1e-6))
            assert len(vals) == m
            step = kids[0, j, 1]
            assert step == tr[i, j, 1] / m
            np.testing.assert_allclose(vals.min() - step / 2, tr[i, j, 0] - tr[i, j, 1] / 2,
                                       atol=1e-9 * tr[i, j, 1])
            np.testing.assert_allclose(vals.max() + step / 2, tr[i, j, 0] + tr[i, j, 1] / 2,
                                       atol=1e-9 * tr[i, j, 1])
        assert np.all(kids[:, P, 0] == tr[i, P, 0]) and np.all(kids[:, P, 1] == 0.0)
        assert np.all(kids[:, -1] == lv[i, -1])


def _small(nsamps=2**17, nseg_ffa=8, seed=4):
    tsamp = 64e-6
    P = 2
    lim = ParamLimits.from_taylor((195.0, 205.0), accel=(-50.0, 50.0)).limits
    mod = {"acc": 12.0}
    cfg = PulsarSearchConfig(nsamps=nsamps, tsamp=tsamp, nbins=32, eta=1.0, param_limits=lim,
                             bseg_brute=nsamps // 256, bseg_ffa=nsamps // nseg_ffa,
                             prune_poly_order=P, ducy_max=0.3, wtsp=1.5, use_fourier=False,
                             branch_max=32)
    sim = PulseSignalConfig(period=1 / 200.37, dt=tsamp, nsamps=nsamps, snr=30.0, ducy=0.1,
                            mod_kwargs=mod)
    ts = sim.generate_analytic(seed=seed)
    fold, pa = orc.ffa_fold(ts.ts_e, ts.ts_v, cfg)
    counts = [len(a) for a in pa]
    dyp_fold = fold.reshape(fold.shape[0], *counts, 2, cfg.nbins)
    pcfg = {"poly_order": P, "nbins": cfg.nbins, "eta": cfg.eta, "branch_max": cfg.branch_max,
            "counts_init": np.array(counts), "param_limits": cfg.param_limits,
            "widths": cfg.score_widths, "tiling": cfg.tiling_strategy}
    return cfg, dyp_fold, pa, pcfg


def test_level_survivors_equal_bruteforce():
    cfg, dyp_fold, pa, pcfg = _small()
    dp = cfg.get_dparams_actual(cfg.niters_ffa)
    res = co.cheby_prune_run(dyp_fold, pa, dp, cfg.tseg_ffa, np.full(7, 3.0), 3, pcfg, 2**20,
                             nlevels=2)
    scheme = res["scheme"]
    c = co.level_coords(scheme, 2)
    tree = res["trees"][0]
    ch, org = co.cheby_branch_batch(tree.leaves, c["cur"], c["prev"], 32, 1.0, 2, 32,
                                    pcfg["tiling"])
    seg = dyp_fold[scheme.get_segment_idx(2)]
    sc = np.empty(len(ch), dtype=np.float32)
    rows, phases = [], []
    for n in range(len(ch)):
        # each child resolved alone, from the series' own derivatives
        s = Chebyshev(ch[n, -2::-1, 0], domain=[c["cur"][0] - c["cur"][1],
                                                 c["cur"][0] + c["cur"][1]])
        ta, ti = c["add"][0], c["init"][0]
        accel = s.deriv(2)(ta)
        vel = s.deriv(1)(ta) - s.deriv(1)(ti)
        freq = ch[n, -1, 0] * (1 - vel / co.C_VAL)
        delay = (s(ta) - s(ti)) / co.C_VAL
        idx, ph = co.cheby_resolve_batch(ch[n:n + 1], c["add"], c["cur"], c["init"],
                                         pcfg["counts_init"], cfg.param_limits, 32)
        np.testing.assert_allclose(
            ph[0], po.phase_idx_vec(ta - ti, ch[n:n + 1, -1, 0], 32, np.array([delay]))[0],
            atol=1e-6)
        ia = po.nearest_2d(np.array([accel]), np.array([freq]), pcfg["counts_init"],
                           cfg.param_limits)[0]
        if not np.array_equal(ia, idx[0]):  # allowed only on a grid-cell edge
            lim = cfg.param_limits
            pos = np.array([(accel - lim[0, 0]) * pcfg["counts_init"][0] / (lim[0, 1] - lim[0, 0]),
                            (freq - lim[1, 0]) * pcfg["counts_init"][1] / (lim[1, 1] - lim[1, 0])])
            assert np.min(np.abs(pos + 1e-6 - np.round(pos + 1e-6))) < 1e-9
        rows.append(seg[idx[0, 0], idx[0, 1]])
        phases.append(ph[0])
        comb = (tree.folds[org[n]] + np.roll(rows[-1], orc.shift_int(ph[0], 32), axis=-1))
        sc[n] = orc.snr_score_batch(comb.astype(np.float32)[None], pcfg["widths"])[0]
    keep = np.where(sc.astype(np.float64) >= 3.0)[0]
    np.testing.assert_array_equal(res["survivors"][1], keep)
    np.testing.assert_array_equal(res["trees"][1].scores, sc[keep])
    exp_leaves = co.cheby_transform_batch(ch[keep], c["next"], c["cur"], pcfg["tiling"])
    np.testing.assert_array_equal(res["trees"][1].leaves, exp_leaves)


def test_branching_pattern_matches_unpruned_counts():
    pa = [np.array([-3.0, 0.0, 3.0]), np.array([150.0, 310.0])]
    dp = np.array([20.0, 1e-3])
    nseg, tseg, ref = 6, 60.0, 2
    bp = co.generate_bp_cheby(pa, dp, tseg, nseg, 32, 1.0, ref, "aggressive")
    pcfg = {"poly_order": 2, "nbins": 32, "eta": 1.0, "branch_max": 4096,
            "counts_init": np.array([3, 2]), "param_limits": np.array([[-5.0, 5.0],
                                                                        [100.0, 400.0]]),
            "widths": np.array([1, 2, 4]), "tiling": "aggressive"}
    prof = np.ones((nseg, 3, 2, 2, 32), dtype=np.float32)
    prof[:, :, :, 0] = RNG.normal(size=(nseg, 3, 2, 32)).astype(np.float32)
    res = co.cheby_prune_run(prof, pa, dp, tseg, np.full(nseg - 1, -np.inf), ref, pcfg, 2**16)
    assert all(s.n_surv < 2**16 for s in res["stats"])
    n_par = [6] + [s.n_surv for s in res["stats"][:-1]]
    got = [s.n_leaves / npar for s, npar in zip(res["stats"], n_par)]
    np.testing.assert_allclose(bp, got, rtol=1e-12)
    assert max(got) > 1.5  # the pattern is not trivial


def test_search_finds_the_taylor_candidate():
    cfg, dyp_fold, pa, pcfg = _small()
    dp = cfg.get_dparams_actual(cfg.niters_ffa)
    thr = np.linspace(2.0, 5.0, 7)
    rt = po.prune_run(dyp_fold, pa, dp, cfg.tseg_ffa, thr, 3, pcfg, 2**16)
    rc = co.cheby_prune_run(dyp_fold, pa, dp, cfg.tseg_ffa, thr, 3, pcfg, 2**16)
    bt = int(np.argmax(rt["tree"].scores))
    bc = int(np.argmax(rc["tree"].scores))
    ft, fc = rt["report"][bt, -3, 0], rc["report"][bc, -3, 0]
    assert abs(fc - 200.37) < 0.02 and abs(fc - ft) < 0.01
    assert abs(rc["report"][bc, 0, 0] - rt["report"][bt, 0, 0]) < 5.0
    st, sc = float(rt["tree"].scores[bt]), float(rc["tree"].scores[bc])
    assert abs(sc - st) / st < 0.1
