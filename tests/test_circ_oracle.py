"""Pin the circular-basis EP oracle (core/circular.py): the circular propagator against
an analytic orbit, the region masks and validation rules on hand-made cells, branching
against the (separately pinned) Taylor branch, the deferred crackle expansion's order
and tiling, resolve against the Taylor resolve and the analytic orbit, the basis-flag
writes, and one level's survivors against a child-by-child brute force.

CPU only; the whole file runs in seconds.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as orc
from oracle import prune_oracle as po
from pyloki_amd.config import ParamLimits, PulsarSearchConfig
from pyloki_amd.simulation.pulse import PulseSignalConfig

RNG = np.random.default_rng(2024)
C = po.C_VAL


def _orbit_derivs(t: float, amp: float, omega: float, phi: float, vsec: float,
                  c0: float) -> np.ndarray:
    """[d5, d4, d3, d2, d1, d0] of d(t) = amp sin(omega t + phi) + vsec t + c0."""
    th = omega * t + phi
    s, c = math.sin(th), math.cos(th)
    return np.array([amp * omega**5 * c, amp * omega**4 * s, -amp * omega**3 * c,
                     -amp * omega**2 * s, amp * omega * c + vsec,
                     amp * s + vsec * t + c0])


@pytest.mark.parametrize("in_hole", [False, True])
def test_circular_shift_matches_analytic_orbit(in_hole):
    for _ in range(50):
        amp = RNG.uniform(1e4, 1e7)
        omega = 2 * np.pi / RNG.uniform(10, 5000)
        phi, vsec, c0 = RNG.uniform(0, 2 * np.pi), RNG.uniform(-3e4, 3e4), RNG.uniform(-1e3, 1e3)
        ti, dt = RNG.uniform(-100, 100), RNG.uniform(-200, 200)
        vi = _orbit_derivs(ti, amp, omega, phi, vsec, c0)
        vj = _orbit_derivs(ti + dt, amp, omega, phi, vsec, c0)
        got = po.shift_taylor_circular(vi[None], dt, in_hole)[0]
        scale = np.array([amp * omega**k for k in (5, 4, 3, 2, 1)] + [amp + abs(vsec * dt)])
        np.testing.assert_allclose(got, vj, rtol=0, atol=1e-9 * scale.max())
        for k in range(5):  # each derivative to its own scale
            assert abs(got[k] - vj[k]) <= 1e-9 * scale[k]


def _cell(crackle, snap, jerk, accel, dsnap=1.0, djerk=1.0, daccel=1.0):
    lv = np.zeros((1, 7, 2))
    lv[0, :4, 0] = [crackle, snap, jerk, accel]
    lv[0, 1:4, 1] = [dsnap, djerk, daccel]
    return lv


def test_circ_mask_regions():
    cases = [  # (cell, expected region)
        (_cell(0.0, -5.0, 0.0, 5.0), "snap"),        # -s*a > 0, both significant
        (_cell(0.0, 5.0, 0.0, 5.0), "taylor"),       # unphysical snap, jerk insignificant
        (_cell(-1.0, 5.0, 3.0, 5.0), "crackle"),     # unphysical snap, -c*j > 0, jerk sig.
        (_cell(1.0, 0.5, 3.0, 5.0), "taylor"),       # hole but -c*j < 0
        (_cell(-1.0, 0.5, 3.0, 5.0), "crackle"),     # snap insignificant, crackle-jerk ok
        (_cell(-1.0, -5.0, 3.0, 1.5), "crackle"),    # accel insignificant
        (_cell(-1.0, -5.0, 1.5, 5.0), "snap"),       # snap region wins over crackle
        (_cell(-1.0, 0.0, 2.0, 5.0), "taylor"),      # |jerk| == 2 djerk is not significant
    ]
    for lv, region in cases:
        ms, mc, mt = po.circ_mask(lv, 2.0)
        assert (bool(ms[0]), bool(mc[0]), bool(mt[0])) == (
            region == "snap", region == "crackle", region == "taylor"), region


def test_circ_validate_rules():
    p_orb_min, xm = 100.0, 2.0e6
    om_max_sq = (2 * np.pi / p_orb_min) ** 2
    rows = []
    # (snap, dsnap, accel, daccel, keep)
    rows.append((-1e-3, 1e-5, 10.0, 0.1, True))              # omega^2 = 1e-4 < max
    rows.append((1e-3, 1e-5, 10.0, 0.1, False))              # significant, unphysical
    rows.append((-100.0 * om_max_sq, 1e-9, 10.0, 0.1, False))  # omega^2 = 10 x max
    a_big = 2.0 * xm * (1e-4) ** (2.0 / 3.0)
    rows.append((-1e-4 * a_big, 1e-9, a_big, 0.1, False))    # |a| above x_mass * w^(4/3)
    rows.append((1e-3, 1.0, 10.0, 0.1, True))                # snap insignificant: hole, kept
    rows.append((1e-3, 1e-5, 10.0, 5.0, True))               # accel insignificant: kept
    lv = np.zeros((len(rows), 7, 2))
    for i, (s, ds, a, da, _) in enumerate(rows):
        lv[i, 1] = [s, ds]
        lv[i, 3] = [a, da]
    kept, org, k = po.circ_validate(lv, np.arange(len(rows)), p_orb_min, xm, 5.0)
    assert k.tolist() == [i for i, r in enumerate(rows) if r[-1]]
    np.testing.assert_array_equal(org, k)


def _leaves5(n: int, coord_cur, crackle_ratio: float) -> np.ndarray:
    """Random 5-parameter cells whose steps are random multiples of the next level's
    steps; the crackle step is crackle_ratio times its next step."""
    nbins, eta = 32, 1.0
    f0 = RNG.uniform(199, 201, n)
    dpf = po.step_f(5, coord_cur[1], nbins, eta)
    dnew = dpf[None, :] * C / f0[:, None]
    lv = np.zeros((n, 7, 2))
    ratios = RNG.uniform(0.3, 4.5, (n, 5))
    ratios[:, 0] = crackle_ratio
    lv[:, :5, 1] = dnew * ratios
    lv[:, :5, 0] = RNG.uniform(-3, 3, (n, 5)) * lv[:, :5, 1]
    lv[:, 5, 0] = RNG.uniform(-100, 100, n)
    lv[:, 6, 0] = f0
    lv[:, 6, 1] = RNG.integers(0, 3, n)
    return lv


def test_circ_branch_equals_taylor_when_crackle_needs_no_branching():
    coord = (3.0, 4.0)
    lv = _leaves5(200, coord, crackle_ratio=0.5)
    got, org = po.circ_branch_batch(lv, coord, 32, 1.0, 64, 2.0)
    exp, eorg = po.branch_batch(lv, coord, 32, 1.0, 5, 64)
    np.testing.assert_array_equal(org, eorg)
    np.testing.assert_array_equal(got, exp)


def test_circ_branch_crackle_expansion_order_and_tiling():
    coord = (3.0, 4.0)
    lv = _leaves5(120, coord, crackle_ratio=3.7)
    # make half of the parents' children land in the hole: snap insignificant, jerk large
    lv[::2, 1, 0] = 0.0
    lv[::2, 2, 0] = 40.0 * lv[::2, 2, 1]
    got, org = po.circ_branch_batch(lv, coord, 32, 1.0, 64, 2.0)
    # stage 1 = the Taylor branch with crackle held: branch a copy whose crackle does not
    # branch, then give every child the crackle step the circular branch assigns
    held = lv.copy()
    held[:, 0, 1] = 1e-300
    st1, sorg = po.branch_batch(held, coord, 32, 1.0, 5, 64)
    dnew0 = po.step_f(5, coord[1], 32, 1.0)[0] * C / lv[:, 6, 0]
    num = np.maximum(1, np.ceil(lv[:, 0, 1] / dnew0 - 1e-6)).astype(int)
    assert np.all(num == 4)
    st1[:, 0, 1] = (lv[:, 0, 1] / num)[sorg]
    sig = 2.0
    s, j, a = st1[:, 1, 0], st1[:, 2, 0], st1[:, 3, 0]
    ds, dj, da = st1[:, 1, 1], st1[:, 2, 1], st1[:, 3, 1]
    m_snap = (np.abs(s) > sig * ds) & (np.abs(a) > sig * da) & (-s * a > 0)
    expand = ~m_snap & (np.abs(j) > sig * dj)
    assert expand.sum() > 10 and (~expand).sum() > 10
    nk = int((~expand).sum())
    assert len(got) == nk + 4 * int(expand.sum())
    np.testing.assert_array_equal(got[:nk], st1[~expand])
    np.testing.assert_array_equal(org[:nk], sorg[~expand])
    exp_rows = st1[expand]
    for e in range(len(exp_rows)):
        blk = got[nk + 4 * e: nk + 4 * (e + 1)]
        p = sorg[expand][e]
        assert np.all(org[nk + 4 * e: nk + 4 * (e + 1)] == p)
        np.testing.assert_array_equal(blk[:, 1:], np.repeat(exp_rows[e:e + 1, 1:], 4, axis=0))
        pc, dc = lv[p, 0, 0], lv[p, 0, 1]
        np.testing.assert_allclose(blk[:, 0, 0], pc - dc / 2 + (np.arange(4) + 0.5) * dc / 4,
                                   rtol=1e-12, atol=1e-12 * dc)
        assert np.all(blk[:, 0, 1] == dc / 4)


def test_circ_resolve_taylor_region_equals_taylor_resolve():
    lv = _leaves5(300, (3.0, 4.0), crackle_ratio=1.0)
    lv[:, 1, 0] = 0.0    # snap insignificant
    lv[:, 2, 0] = 0.0    # jerk insignificant -> every cell is Taylor
    lim = np.array([[-1e6, 1e6]] * 4 + [[198.0, 202.0]])
    counts = np.array([1, 1, 1, 41, 801])
    args = ((12.0, 1.0), (3.0, 4.0), (0.0, 1.0), counts, lim, 32)
    exp = po.resolve_batch(lv, *args)
    got = po.circ_resolve_batch(lv.copy(), *args, 2.0)
    np.testing.assert_array_equal(got[0], exp[0])
    np.testing.assert_array_equal(got[1], exp[1])


def test_circ_resolve_values_follow_the_orbit():
    """A snap-region cell at t_cur resolves to the orbit's accel at t_add, its frequency
    shifted by the velocity gained since t_init, and the delay difference."""
    t_cur, t_add, t_init = 7.0, 19.0, 0.0
    for _ in range(30):
        amp, omega = RNG.uniform(1e5, 1e6), 2 * np.pi / RNG.uniform(30, 300)
        phi, vsec = RNG.uniform(0, 2 * np.pi), RNG.uniform(-1e3, 1e3)
        d = _orbit_derivs(t_cur, amp, omega, phi, vsec, 0.0)
        if abs(d[1]) < 1e-3 * amp * omega**4 or abs(d[3]) < 1e-3 * amp * omega**2:
            continue
        lv = np.zeros((1, 7, 2))
        lv[0, :6, 0] = d
        lv[0, :4, 1] = 1e-12     # tiny errors: every quantity significant
        lv[0, 6, 0] = 200.0
        acc, freq, delay, masks = po.circ_resolve_values(lv, (t_add, 1.0), (t_cur, 1.0),
                                                         (t_init, 1.0), 2.0)
        assert masks[0][0]
        da = _orbit_derivs(t_add, amp, omega, phi, vsec, 0.0)
        di = _orbit_derivs(t_init, amp, omega, phi, vsec, 0.0)
        assert acc[0] == pytest.approx(da[3], rel=1e-9, abs=1e-9 * amp * omega**2)
        assert freq[0] == pytest.approx(200.0 * (1 - (da[4] - di[4]) / C), rel=1e-13)
        assert delay[0] == pytest.approx((da[5] - di[5]) / C, rel=1e-9, abs=1e-12)


def test_circ_flag_writes():
    lim = np.array([[-1e6, 1e6]] * 4 + [[198.0, 202.0]])
    counts = np.array([1, 1, 1, 41, 801])
    snap = _cell(0.0, -5.0, 0.0, 5.0)
    crack = _cell(-1.0, 0.5, 3.0, 5.0)
    tayl = _cell(0.0, 0.0, 0.0, 0.0)
    for lv in (snap, crack, tayl):
        lv[0, 6] = [200.0, 7.0]  # flag 7 marks "untouched"
    args = ((12.0, 1.0), (3.0, 4.0), (0.0, 1.0), counts, lim, 32, 2.0)
    batch = np.concatenate([snap, crack, tayl])
    po.circ_resolve_batch(batch, *args)
    assert batch[:, 6, 1].tolist() == [1.0, 0.0, 7.0]   # Taylor present: crackle -> 0
    batch = np.concatenate([snap, crack])
    po.circ_resolve_batch(batch, *args)
    assert batch[:, 6, 1].tolist() == [1.0, 2.0]


def _synthetic_level(n_par: int = 48, seed: int = 5):
    """Parents at level 6 of a 16-segment middle-out scheme whose children cover every
    branch of the circular logic: snap-circular cells near an orbit, hole cells (snap
    insignificant, jerk significant) whose crackle must branch, Taylor cells, cells that
    validation drops; random parent folds and segment rows."""
    rng = np.random.default_rng(seed)
    scheme = po.MiddleOut(16, 7, 0.5)
    L = 6
    coords = {"init": scheme.get_coord(0), "cur": scheme.get_current_coord(L),
              "next": scheme.get_coord(L), "add": scheme.get_segment_coord(L)}
    nbins = 32
    f0 = rng.uniform(199.5, 200.5, n_par)
    dnew = po.step_f(5, coords["cur"][1], nbins, 1.0)[None, :] * C / f0[:, None]
    lv = np.zeros((n_par, 7, 2))
    lv[:, :5, 1] = dnew * rng.uniform(0.6, 2.6, (n_par, 5))
    lv[:, 0, 1] = dnew[:, 0] * 3.3                      # crackle always needs branching
    kind = np.arange(n_par) % 4
    omega = 2 * np.pi / 40.0
    amp = 30.0 * lv[:, 3, 1] / omega**2                 # accel amplitude 30 accel cells
    th = rng.uniform(0, 2 * np.pi, n_par)
    snapc = kind == 0                                   # on a circular orbit
    lv[snapc, 3, 0] = -amp[snapc] * omega**2 * np.sin(th[snapc])
    lv[snapc, 2, 0] = -amp[snapc] * omega**3 * np.cos(th[snapc])
    lv[snapc, 1, 0] = amp[snapc] * omega**4 * np.sin(th[snapc])
    lv[snapc, 1, 1] = np.abs(lv[snapc, 1, 0]) / 40.0    # snap well measured
    hole = kind == 1                                    # snap ~ 0, jerk significant
    lv[hole, 1, 0] = 0.0
    lv[hole, 2, 0] = 25.0 * lv[hole, 2, 1] * np.sign(rng.normal(size=hole.sum()))
    lv[hole, 3, 0] = rng.uniform(-5, 5, hole.sum()) * lv[hole, 3, 1]
    bad = kind == 2                                     # significant but unphysical snap
    lv[bad, 3, 0] = 40.0 * lv[bad, 3, 1]
    lv[bad, 1, 0] = 40.0 * lv[bad, 1, 1]
    # kind 3: Taylor (everything near zero)
    lv[:, 4, 0] = rng.uniform(-2, 2, n_par) * lv[:, 4, 1]
    lv[:, 5, 0] = rng.uniform(-50, 50, n_par)
    lv[:, 6, 0] = f0
    lv[:, 6, 1] = rng.integers(0, 3, n_par)
    folds = np.abs(rng.normal(size=(n_par, 2, nbins))).astype(np.float32)
    folds[:, 1] += 1.0
    tree = po.Tree(lv, folds, orc.snr_score_batch(folds, np.array([1, 2, 4])),
                   np.zeros((n_par, 7), np.int32))
    n_acc, n_freq = 64, 512
    seg = np.abs(rng.normal(size=(1, 1, 1, n_acc, n_freq, 2, nbins))).astype(np.float32)
    seg[..., 1, :] += 1.0
    lim = np.array([[-1e3, 1e3], [-1e3, 1e3], [-1e4, 1e4], [-3e3, 3e3], [199.0, 201.0]])
    pcfg = {"poly_order": 5, "nbins": nbins, "eta": 1.0, "branch_max": 64,
            "counts_init": np.array([1, 1, 1, n_acc, n_freq]), "param_limits": lim,
            "widths": np.array([1, 2, 4]), "tiling": "aggressive", "prop_sig": 2.0,
            "val_sig": 5.0, "p_orb_min": 10.0, "x_mass_const": 1.0e8}
    return tree, seg, coords, pcfg


def test_circ_level_survivors_equal_bruteforce():
    """One level on synthetic parents: the children of branch + validate are resolved,
    folded and scored one by one; survivors are the children at or above the threshold
    in child order, carrying the basis flags and the transformed parameters."""
    tree, seg, coords, pcfg = _synthetic_level()
    ch, org = po.circ_branch_batch(tree.leaves, coords["cur"], 32, 1.0, 64, 2.0)
    n_branch = len(ch)
    nexp = int(np.sum(ch[:, 0, 0] != tree.leaves[org, 0, 0]))
    ch, org, _ = po.circ_validate(ch, org, pcfg["p_orb_min"], pcfg["x_mass_const"], 5.0)
    assert nexp > 0 and len(ch) < n_branch
    ms, mc, mt = po.circ_mask(ch, 2.0)
    assert ms.any() and mc.any() and mt.any()
    sc = np.empty(len(ch), dtype=np.float32)
    rows = np.empty((len(ch), 2), dtype=np.int64)
    phases = np.empty(len(ch))
    for n in range(len(ch)):
        idx, ph = po.circ_resolve_batch(ch[n:n + 1].copy(), coords["add"], coords["cur"],
                                        coords["init"], pcfg["counts_init"],
                                        pcfg["param_limits"], 32, 2.0)
        rows[n], phases[n] = idx[0, -2:], ph[0]
        row = seg[0, 0, 0, idx[0, -2], idx[0, -1]]
        comb = (tree.folds[org[n]] + np.roll(row, orc.shift_int(ph[0], 32), axis=-1))
        sc[n] = orc.snr_score_batch(comb.astype(np.float32)[None], pcfg["widths"])[0]
    thr = float(np.median(sc))
    new, st, surv = po.circ_prune_level(tree, seg, coords, thr, pcfg, max_sugg=2**20)
    keep = np.where(sc.astype(np.float64) >= thr)[0]
    np.testing.assert_array_equal(surv, keep)
    np.testing.assert_array_equal(new.scores, sc[keep])
    assert (st.n_leaves, st.n_leaves_phy, st.n_surv) == (n_branch, len(ch), len(keep))
    # flags: snap 1; crackle 0 (the batch holds Taylor cells); Taylor: the parent's
    flags = ch[:, 6, 1].copy()
    flags[ms] = 1.0
    flags[mc] = 0.0
    np.testing.assert_array_equal(new.leaves[:, 6, 1], flags[keep])
    np.testing.assert_array_equal(new.backtracks[:, 0], org[keep])
    np.testing.assert_array_equal(new.backtracks[:, 4:6], rows[keep])
    np.testing.assert_array_equal(new.backtracks[:, -1], phases[keep].astype(np.int32))
    # parameters carried to the next coordinate by the region's propagator
    dt = coords["next"][0] - coords["cur"][0]
    for n_out, n in enumerate(keep):
        vec = ch[n, :6, 0][None]
        if ms[n]:
            exp = po.shift_taylor_circular(vec, dt, False)[0]
        elif mc[n]:
            exp = po.shift_taylor_circular(vec, dt, True)[0]
        else:
            exp = po.shift_taylor(vec, dt)[0]
        np.testing.assert_array_equal(new.leaves[n_out, :6, 0], exp)
