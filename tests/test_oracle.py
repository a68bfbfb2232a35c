"""Pin the CPU oracle against what the mathematics fixes, independent of its own formulas.

Each check is chosen so that a plausible restatement mistake (a dropped term, a wrong sign,
an off-by-one index, a transposed operand, a different rounding) fails it. CPU only.
"""

from __future__ import annotations

import math
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle import oracle as orc
from pyloki_amd.config import PulsarSearchConfig
from pyloki_amd.utils.misc import C_VAL

RNG = np.random.default_rng(1234)


def exact_bin(t: float, f: float, nbins: int, delay: float = 0.0):
    """Exact rational phase bin (half-up) and its distance to a rounding boundary."""
    x = (Fr(t) - Fr(delay)) * Fr(f)
    frac = x - math.floor(x)
    pos = frac * nbins
    b = math.floor(pos + Fr(1, 2))
    dist = min(abs(pos + Fr(1, 2) - math.floor(pos + Fr(1, 2))),
               abs(pos + Fr(1, 2) - math.floor(pos + Fr(1, 2)) - 1))
    return b % nbins, float(dist), float(pos)


def test_phase_idx_int_vs_exact_rational():
    n_checked = 0
    for _ in range(4000):
        t = float(RNG.uniform(-50, 50))
        f = float(RNG.uniform(0.5, 700))
        nbins = int(RNG.choice([32, 64, 128, 100]))
        delay = float(RNG.uniform(-1e-3, 1e-3))
        b, dist, pos = exact_bin(t, f, nbins, delay)
        if dist < 1e-6:
            continue
        assert orc.phase_idx_int(t, f, nbins, delay) == b
        assert abs(orc.phase_idx(t, f, nbins, delay) - pos) < 1e-6 * nbins or pos > nbins - 1e-6
        n_checked += 1
    assert n_checked > 3900


def test_phase_idx_special_cases():
    # An exact integer number of cycles lands on bin 0; a half cycle on bin nbins/2.
    assert orc.phase_idx(2.0, 3.0, 64, 0.0) == 0.0
    assert orc.phase_idx(0.5, 1.0, 64, 0.0) == 32.0
    assert orc.phase_idx_int(0.5, 1.0, 64, 0.0) == 32
    # Just below a full cycle rounds up to nbins, which wraps to 0.
    assert orc.phase_idx_int(1.0 - 1e-9, 1.0, 64, 0.0) == 0
    # Negative times wrap into [0, nbins).
    assert orc.phase_idx_int(-0.25, 1.0, 64, 0.0) == 48
    assert orc.shift_int(63.4999, 64) == 63 and orc.shift_int(63.5, 64) == 0


def _powi_chain(a: float, b: int) -> float:
    # compiler-rt __powidf2, transcribed
    r = 1.0
    while True:
        if b & 1:
            r = r * a
        b = b // 2
        if b == 0:
            return r
        a = a * a


def test_powi_is_the_binary_powering_chain():
    xs = np.concatenate([RNG.uniform(-3000, 3000, 1000), RNG.uniform(-1, 1, 1000)])
    for p in range(7):
        for x in xs:
            got = orc.powi(float(x), p)
            assert got == _powi_chain(float(x), p)
            exact = Fr(float(x)) ** p
            if exact != 0:
                assert abs(Fr(got) - exact) <= abs(exact) * p * Fr(2) ** -52


def test_shift_taylor_params_closed_form():
    # [d_3, d_2, d_1, d_0] shifted by dt: d_m(dt) = sum_k d_k dt^(k-m)/(k-m)!.
    for _ in range(200):
        vec = RNG.normal(size=4) * np.array([1.0, 100.0, 1e4, 1e6])
        dt = float(RNG.uniform(-500, 500))
        out = orc.shift_taylor_params(vec, dt)
        k = [3, 2, 1, 0]
        for i, m in enumerate(k):
            exact = sum(Fr(vec[j]) * Fr(dt) ** (k[j] - m) / math.factorial(k[j] - m)
                        for j in range(4) if k[j] >= m)
            assert out[i] == pytest.approx(float(exact), rel=1e-12, abs=1e-9)
    # group property: shift(a) then shift(b) == shift(a + b)
    vec = np.array([0.5, -20.0, 3.0, 0.0])
    ab = orc.shift_taylor_params(orc.shift_taylor_params(vec, 3.0), 4.5)
    np.testing.assert_allclose(ab, orc.shift_taylor_params(vec, 7.5), rtol=1e-13, atol=1e-9)


def _exact_resolve(pset, counts_prev, limits, level, latter, tseg_brute, nbins):
    """Exact-rational restatement of ffa_taylor_resolve (fold.py:617-671), with the
    distance of every truncation/rounding to its boundary."""
    P = len(pset)
    if P == 1:
        dt = Fr(latter * 2 ** (level - 1)) * Fr(tseg_brute)
        pnew = [Fr(pset[0])]
        delay = Fr(0)
    else:
        dt = (Fr(latter) - Fr(1, 2)) * 2 ** (level - 1) * Fr(tseg_brute)
        d = [Fr(x) for x in pset[:-1]]  # [d_{P-1}, ..., d_2]
        ks = list(range(P, 1, -1))      # derivative order of each entry
        def deriv(m):
            return sum(dk * dt ** (k - m) / math.factorial(k - m) for dk, k in zip(d, ks) if k >= m)
        pnew = [deriv(k) for k in ks]
        pnew.append(Fr(pset[-1]) * (1 - deriv(1) / Fr(C_VAL)))
        delay = deriv(0) / Fr(C_VAL)
    x = (dt - delay) * Fr(pset[-1])
    pos = (x - math.floor(x)) * nbins
    shift = math.floor(pos + Fr(1, 2)) % nbins
    dist = abs(pos + Fr(1, 2) - round(pos + Fr(1, 2)))
    idx = []
    for k in range(P):
        n = int(counts_prev[k])
        if n <= 1:
            idx.append(0)
            continue
        lo, hi = Fr(limits[k][0]), Fr(limits[k][1])
        raw = n * (pnew[k] - lo) / (hi - lo) + Fr(1, 10**6)
        dist = min(dist, abs(raw - round(raw)))
        idx.append(min(max(math.floor(raw), 0), n - 1))
    return idx, shift, float(dist)


@pytest.mark.parametrize("P", [1, 2, 3])
def test_ffa_resolve_vs_exact_rational(P):
    nbins = 64
    tseg_brute = 0.262144
    if P == 1:
        limits = np.array([[195.0, 205.0]])
    elif P == 2:
        limits = np.array([[-50.0, 50.0], [195.0, 205.0]])
    else:
        limits = np.array([[-2.0, 2.0], [-30.0, 30.0], [495.0, 505.0]])
    checked = 0
    for _ in range(400):
        level = int(RNG.integers(1, 10))
        counts_prev = np.array([int(RNG.integers(1, 40)) for _ in range(P - 1)]
                               + [int(RNG.integers(50, 5000))])
        pset = np.array([RNG.uniform(lo, hi) for lo, hi in limits])
        latter = int(RNG.integers(0, 2))
        idx, ph = orc.ffa_resolve(pset, counts_prev, limits, level, latter, tseg_brute, nbins)
        eidx, eshift, dist = _exact_resolve(pset, counts_prev, limits, level, latter,
                                            tseg_brute, nbins)
        if dist < 1e-6:
            continue
        assert idx.tolist() == eidx
        assert orc.shift_int(ph, nbins) == eshift
        checked += 1
    assert checked > 380


def _py_brutefold(ts_e, ts_v, freqs, seg_len, nbins, tsamp, t_ref):
    """Pure-Python loop with exact-rational bins (tiny inputs only)."""
    nseg = -(-len(ts_e) // seg_len)
    out = np.zeros((nseg, len(freqs), 2, nbins), dtype=np.float32)
    for s in range(nseg):
        for fi, f in enumerate(freqs):
            for i in range(min(seg_len, len(ts_e) - s * seg_len)):
                pt = i * tsamp - t_ref
                b, dist, _ = exact_bin(pt, f, nbins)
                assert dist > 1e-9
                out[s, fi, 0, b] = np.float32(out[s, fi, 0, b] + ts_e[s * seg_len + i])
                out[s, fi, 1, b] = np.float32(out[s, fi, 1, b] + ts_v[s * seg_len + i])
    return out


def test_brutefold_vs_python_loop_and_invariants():
    nbins, seg_len, tsamp = 16, 96, 0.013
    ts_e = RNG.normal(size=4 * seg_len).astype(np.float32)
    ts_v = RNG.uniform(0.5, 2, size=4 * seg_len).astype(np.float32)
    freqs = np.array([0.917, 1.7, 3.31])
    for t_ref in (0.0, seg_len * tsamp / 2):
        got = orc.brutefold(ts_e, ts_v, freqs, seg_len, nbins, tsamp, t_ref)
        exp = _py_brutefold(ts_e, ts_v, freqs, seg_len, nbins, tsamp, t_ref)
        np.testing.assert_array_equal(got, exp)
    # Integer-valued data: sums are exact in any order -> conservation is exact.
    ti = RNG.integers(-50, 50, size=64 * 256).astype(np.float32)
    ones = np.ones_like(ti)
    fold = orc.brutefold(ti, ones, np.array([1.3, 2.9, 7.7]), 256, 64, 1e-3, 0.0)
    seg_sums = ti.reshape(64, 256).sum(axis=1)
    np.testing.assert_array_equal(fold[:, :, 0, :].sum(axis=-1), np.repeat(seg_sums[:, None], 3, 1))
    np.testing.assert_array_equal(fold[:, :, 1, :].sum(axis=-1), np.full((64, 3), 256.0))


def _ffa_cfg(P: int, nsamps: int = 2**14, nbins: int = 32) -> PulsarSearchConfig:
    if P == 1:
        lim = np.array([[3.0, 5.0]])
    else:
        lim = np.array([[-300.0, 300.0], [30.0, 34.0]])
    return PulsarSearchConfig(nsamps=nsamps, tsamp=2e-3, nbins=nbins, eta=1.0,
                              param_limits=lim, bseg_brute=256, use_fourier=False)


@pytest.mark.parametrize("P", [1, 2])
def test_unify_fold_tree_replay(P):
    """Level-L rows equal the sum of level-0 rows rolled by the composed resolved shifts."""
    cfg = _ffa_cfg(P)
    ts_e = RNG.integers(-20, 20, cfg.nsamps).astype(np.float32)
    ts_v = np.ones(cfg.nsamps, dtype=np.float32)
    pa0 = cfg.get_param_arr(cfg.get_dparams(0))
    t_ref = 0.0 if P == 1 else cfg.bseg_brute * cfg.tsamp / 2
    folds = [orc.brutefold(ts_e, ts_v, pa0[-1], cfg.bseg_brute, cfg.nbins, cfg.tsamp, t_ref)]
    tables = []
    counts_prev = cfg.get_param_grid_count(0)
    for level in range(1, cfg.niters_ffa + 1):
        pa = cfg.get_param_arr(cfg.get_dparams(level))
        out, rows, phases = orc.unify_fold(folds[-1], counts_prev, pa, cfg.param_limits, level,
                                           cfg.tseg_brute, cfg.nbins, return_table=True)
        shifts = np.vectorize(lambda x: orc.shift_int(x, cfg.nbins))(phases)
        folds.append(out)
        tables.append((rows, shifts))
        counts_prev = cfg.get_param_grid_count(level)

    def replay(level, q, j):
        if level == 0:
            return folds[0][q, j].astype(np.float64)
        rows, shifts = tables[level - 1]
        t = replay(level - 1, 2 * q, rows[j, 0])
        h = replay(level - 1, 2 * q + 1, rows[j, 1])
        return np.roll(t, shifts[j, 0], axis=-1) + np.roll(h, shifts[j, 1], axis=-1)

    last = len(folds) - 1
    for j in RNG.integers(0, folds[last].shape[1], 12):
        np.testing.assert_array_equal(folds[last][0, j], replay(last, 0, int(j)))
    # per-level conservation: every row keeps the full e-sum of its segment pair span
    for level in range(1, last + 1):
        nseg = folds[level].shape[0]
        span = ts_e.reshape(nseg, -1).sum(axis=1)
        np.testing.assert_array_equal(folds[level][:, :, 0, :].sum(-1),
                                      np.repeat(span[:, None], folds[level].shape[1], 1))


def test_ffa_recovers_injected_frequency():
    """P=1 FFA peak lands at the injected frequency (FFA approximates the direct fold)."""
    from pyloki_amd.simulation.pulse import PulseSignalConfig

    cfg = PulsarSearchConfig(nsamps=2**16, tsamp=1e-3, nbins=32, eta=1.0,
                             param_limits=np.array([[3.0, 5.0]]), bseg_brute=512,
                             ducy_max=0.3, wtsp=1.5, use_fourier=False)
    sim = PulseSignalConfig(period=1 / 4.1234, dt=1e-3, nsamps=2**16, snr=40, ducy=0.08)
    ts = sim.generate_analytic(seed=3)
    snr = orc.ffa_search_snr(ts.ts_e, ts.ts_v, cfg)
    freqs = cfg.get_param_arr(cfg.get_dparams(cfg.niters_ffa))[0]
    fbest = freqs[np.unravel_index(np.argmax(snr), snr.shape)[0]]
    assert abs(fbest - 4.1234) < 2 * (freqs[1] - freqs[0])
    assert snr.max() > 25


def _bf_matched_filter(x: np.ndarray, w: int) -> float:
    """Max over circular shifts of the zero-mean unit-energy boxcar template (FP64)."""
    B = len(x)
    t = np.zeros(B)
    t[:w] = 1.0
    t -= t.mean()
    t /= np.sqrt(np.dot(t, t))
    return max(float(np.dot(np.roll(t, s), x)) for s in range(B))


def test_boxcar_snr_matches_matched_filter():
    widths = np.array([1, 2, 3, 5, 8, 13, 21], dtype=np.int64)
    for B in (32, 64):
        x = RNG.normal(size=(20, B)).astype(np.float32)
        got = orc.boxcar_snr(x, widths)
        for p in range(20):
            for k, w in enumerate(widths):
                assert got[p, k] == pytest.approx(_bf_matched_filter(x[p].astype(np.float64), int(w)),
                                                  rel=2e-5, abs=2e-5)
        ev = np.stack([x * 2.0, np.full_like(x, 4.0)], axis=1)  # e/sqrt(v) == x exactly
        np.testing.assert_array_equal(orc.snr_score_batch(ev, widths), got.max(axis=1))


def test_boxcar_snr_closed_forms():
    B = 64
    widths = np.array([1, 4, 16, 32], dtype=np.int64)
    imp = np.zeros((1, B), dtype=np.float32)
    imp[0, 17] = 1.0
    got = orc.boxcar_snr(imp, widths)[0]
    for k, w in enumerate(widths):
        assert got[k] == pytest.approx(math.sqrt((B - w) / (B * w)), rel=1e-6)
    # two-level profile: ones on a block of width W -> S/N at width W = sqrt(W (B-W) / B)
    for W in (4, 16):
        prof = np.zeros((1, B), dtype=np.float32)
        prof[0, 60:] = 1.0
        prof[0, : W - 4] = 1.0
        s = orc.boxcar_snr(prof, np.array([W], dtype=np.int64))[0, 0]
        assert s == pytest.approx(math.sqrt(W * (B - W) / B), rel=1e-6)
    # constant profile has no signal: S/N == 0 for every width
    const = np.full((1, B), 3.0, dtype=np.float32)
    np.testing.assert_allclose(orc.boxcar_snr(const, widths)[0], 0.0, atol=1e-5)
