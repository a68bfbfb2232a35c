"""Independent pins for oracle functions that a restatement alone would leave unpinned:

- the EP report (Taylor, circular and Chebyshev bases) against the exact rational gauge
  transform, with line-of-sight velocities large enough that a wrong sign or a missing
  factor of the Doppler term fails;
- the ascend composition on integer-valued folds against the sum over segments of each
  segment row rolled by its shift, the row index and shift taken from an exact rational
  resolve;
- the quadrature and conservative width propagation of the Taylor transform and of the
  Chebyshev domain transfer against their closed forms;
- the BF16 rounding against torch's round-to-nearest-even over every exponent and
  mantissa around the ties;
- the FP64 FFA S/N against a 50-digit decimal evaluation on integer data (the folds are
  then exact in both FP32 and FP64).

CPU only.
"""

from __future__ import annotations

import math
from decimal import Decimal, getcontext
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle import cheby_oracle as cb
from oracle import oracle as orc
from oracle import prune_oracle as po
from pyloki_amd.config import ParamLimits, PulsarSearchConfig

RNG = np.random.default_rng(2024)
C = Fr(po.C_VAL)


def _fsqrt(x: Fr) -> Fr:
    """Square root of a non-negative rational, correct far beyond FP64 (60 digits)."""
    getcontext().prec = 60
    return Fr(Decimal(x.numerator) / Decimal(x.denominator)).limit_denominator(10**40) \
        if x == 0 else Fr((Decimal(x.numerator) / Decimal(x.denominator)).sqrt())


def _gauge_exact(vals, sigs, v, dv, f0):
    s = 1 - v / C
    out_v = [x / s for x in vals]
    out_s = [_fsqrt((sg / s) ** 2 + (x / (C * s * s)) ** 2 * dv * dv)
             for x, sg in zip(vals, sigs)]
    return out_v, out_s, f0 * s, f0 * dv / C


def _taylor_leaves(n: int, P: int) -> np.ndarray:
    lv = np.zeros((n, P + 2, 2))
    for k in range(P - 1):  # the derivatives above the velocity
        lv[:, k, 0] = RNG.uniform(-1, 1, n) * 10.0 ** RNG.integers(-6, 3, n)
        lv[:, k, 1] = np.abs(lv[:, k, 0]) * RNG.uniform(1e-3, 1e-1, n)
    lv[:, -3, 0] = RNG.uniform(-6e6, 6e6, n)   # v: up to 2 % of c
    lv[:, -3, 1] = RNG.uniform(1e3, 3e5, n)
    lv[:, -2, 0] = RNG.uniform(-1e3, 1e3, n)
    lv[:, -1, 0] = RNG.uniform(1.0, 700.0, n)
    lv[:, -1, 1] = RNG.integers(0, 3, n)
    return lv


@pytest.mark.parametrize("P", [2, 3, 5])
def test_report_gauge_transform_exact(P):
    """poly_taylor_report_batch (core/taylor.py:342-360), also the circular basis's report
    (dyn_circular_taylor.py:726-740)."""
    lv = _taylor_leaves(200, P)
    got = po.report_batch(lv)
    worst = 0.0
    for n in range(len(lv)):
        vals = [Fr(x) for x in lv[n, :-3, 0]]
        sigs = [Fr(x) for x in lv[n, :-3, 1]]
        ov, os_, f, df = _gauge_exact(vals, sigs, Fr(lv[n, -3, 0]), Fr(lv[n, -3, 1]),
                                      Fr(lv[n, -1, 0]))
        for k in range(P - 1):
            for g, e in ((got[n, k, 0], ov[k]), (got[n, k, 1], os_[k])):
                worst = max(worst, abs(float(Fr(g) - e)) / max(abs(float(e)), 1e-300))
        worst = max(worst, abs(float(Fr(got[n, -3, 0]) - f)) / float(f))
        worst = max(worst, abs(float(Fr(got[n, -3, 1]) - df)) / float(df))
        np.testing.assert_array_equal(got[n, -2:], lv[n, -2:])
    assert worst <= 4e-16
    # The Doppler term is large here: dropping it or flipping its sign is a ~1 % error.
    s = 1 - lv[:, -3, 0] / po.C_VAL
    assert np.max(np.abs(s - 1)) > 1e-2
    wrong = lv[:, :-3, 0] * s[:, None]
    assert np.max(np.abs(wrong - got[:, :-3, 0]) / np.abs(got[:, :-3, 0])) > 1e-2


def _cheb_poly(n: int) -> list[int]:
    """Integer power-basis coefficients of T_n (ascending powers)."""
    t0, t1 = [1], [0, 1]
    if n == 0:
        return t0
    for _ in range(n - 1):
        nxt = [0] * (len(t1) + 1)
        for i, c in enumerate(t1):
            nxt[i + 1] += 2 * c
        for i, c in enumerate(t0):
            nxt[i] -= c
        t0, t1 = t1, nxt
    return t1


@pytest.mark.parametrize("P", [2, 3, 4])
def test_cheby_report_exact(P):
    """poly_chebyshev_report_batch (core/chebyshev.py:371-398): the Chebyshev series on the
    domain (t0, ts) converted to Taylor derivatives at t0 (d_k = k! c_k / ts^k, c_k the
    x^k coefficient of sum a_n T_n(x)), errors in quadrature through the same map, then
    the gauge transform."""
    K = P + 1  # coefficients a_K .. a_0 (rows [:-1])
    n = 60
    lv = np.zeros((n, K + 2, 2))
    lv[:, :-1, 0] = RNG.uniform(-1, 1, (n, K + 1)) * 10.0 ** RNG.integers(0, 4, (n, K + 1))
    lv[:, -3, 0] = RNG.uniform(2e3, 8e4, n) * RNG.choice([-1, 1], n)  # a_1 sets v
    lv[:, :-1, 1] = np.abs(lv[:, :-1, 0]) * RNG.uniform(1e-4, 1e-2, (n, K + 1))
    lv[:, -1, 0] = RNG.uniform(50.0, 700.0, n)
    ts = 37.5
    got = cb.cheby_report_batch(lv, (0.0, ts))
    T = [_cheb_poly(m) for m in range(K + 1)]
    for i in range(n):
        a = [Fr(x) for x in lv[i, :-1, 0][::-1]]   # a_0 .. a_K
        da = [Fr(x) for x in lv[i, :-1, 1][::-1]]
        d, dd, mag = [], [], []
        for k in range(K + 1):
            f = Fr(math.factorial(k)) / Fr(ts) ** k
            ck = sum((a[m] * T[m][k] for m in range(k, K + 1) if k < len(T[m])), Fr(0))
            vk = sum((da[m] ** 2 * T[m][k] ** 2 for m in range(k, K + 1) if k < len(T[m])),
                     Fr(0))
            mk = sum((abs(a[m] * T[m][k]) for m in range(k, K + 1) if k < len(T[m])), Fr(0))
            d.append(ck * f)
            dd.append(_fsqrt(vk) * f)
            mag.append(mk * f)
        # d[0] delay, d[1] velocity, d[2..K] the derivatives above it
        np.testing.assert_allclose(got[i, -2, 0], float(d[0]), rtol=0,
                                   atol=1e-14 * float(mag[0]))
        vals = [d[k] for k in range(K, 1, -1)]
        sigs = [dd[k] for k in range(K, 1, -1)]
        ov, os_, f, df = _gauge_exact(vals, sigs, d[1], dd[1], Fr(lv[i, -1, 0]))
        for j in range(K - 1):
            k = K - j
            tol = 1e-14 * float(mag[k] / abs(1 - d[1] / C))
            assert abs(got[i, j, 0] - float(ov[j])) <= tol, (i, j)
            assert abs(got[i, j, 1] - float(os_[j])) <= 1e-13 * float(os_[j]) + 1e-300, (i, j)
        assert abs(got[i, -3, 0] - float(f)) <= 1e-14 * float(f) + 1e-14 * float(
            Fr(lv[i, -1, 0]) * mag[1] / C)
        assert abs(got[i, -3, 1] - float(df)) <= 1e-13 * float(df)


def _tmat(dt: Fr, n: int):
    return [[dt ** (i - j) / math.factorial(i - j) if j <= i else Fr(0) for j in range(n)]
            for i in range(n)]


@pytest.mark.parametrize("tiling", ["aggressive", "conservative", "quadrature"])
def test_transform_width_propagation_closed_form(tiling):
    """shift_taylor_full (utils/transforms.py:93-137): values v' = T v; widths by the tiling
    strategy: |diag T| e, |T| e, or sqrt(T^2 e^2), T[i][j] = dt^(i-j)/(i-j)!."""
    P = 4
    lv = _taylor_leaves(50, P)
    lv[:, :-1, 1] = RNG.uniform(0.1, 10.0, (50, P + 1))
    dt = -173.25  # negative: the absolute values matter
    out = po.transform_batch(lv, (dt, 1.0), (0.0, 1.0), tiling)
    T = _tmat(Fr(dt), P + 1)
    for n in range(len(lv)):
        e = [Fr(x) for x in lv[n, :-1, 1]]
        for i in range(P + 1):
            if tiling == "aggressive":
                exp = abs(T[i][i]) * e[i]
            elif tiling == "conservative":
                exp = sum((abs(T[i][j]) * e[j] for j in range(P + 1)), Fr(0))
            else:
                exp = _fsqrt(sum((T[i][j] ** 2 * e[j] ** 2 for j in range(P + 1)), Fr(0)))
            assert abs(out[n, i, 1] - float(exp)) <= 1e-14 * float(exp), (tiling, n, i)
    np.testing.assert_array_equal(out[:, -1], lv[:, -1])
    # The strategies differ by far more than the tolerance on these cells.
    agg = po.transform_batch(lv, (dt, 1.0), (0.0, 1.0), "aggressive")[:, -2, 1]
    if tiling != "aggressive":
        assert np.min(out[:, -2, 1] / agg) > 10.0


@pytest.mark.parametrize("tiling", ["conservative", "quadrature"])
def test_cheby_tile_errors_closed_form(tiling):
    """The Chebyshev domain transfer's widths (utils/transforms.py shift_cheby_errors):
    |C|^T e or sqrt((C^2)^T e^2), with C the transfer matrix (coefficients move as
    a' = a C)."""
    K = 4
    Cm = RNG.normal(size=(K + 1, K + 1))
    e = RNG.uniform(0.1, 2.0, (20, K + 1))
    got = cb._tile_errors(e, Cm, tiling)
    for n in range(len(e)):
        for k in range(K + 1):
            if tiling == "conservative":
                exp = sum(Fr(e[n, j]) * abs(Fr(Cm[j, k])) for j in range(K + 1))
            else:
                exp = _fsqrt(sum(Fr(e[n, j]) ** 2 * Fr(Cm[j, k]) ** 2 for j in range(K + 1)))
            assert abs(got[n, k] - float(exp)) <= 1e-14 * float(exp)


def test_ascend_composition_integer_folds():
    """ascend (dyn_poly_taylor.py:587-628 with taylor.py:280-320 and common.py:432-459) on
    integer-valued segment folds: every survivor's fold is the sum over the segments so
    far of its segment row rolled by its shift, row and shift from an exact rational
    resolve at each segment's reference time (exact sums in FP32)."""
    B, P = 32, 2
    nseg, n_acc, n_freq = 8, 5, 41
    limits = np.array([[-20.0, 20.0], [199.0, 201.0]])
    dyp_fold = RNG.integers(-50, 50, (nseg, n_acc, n_freq, 2, B)).astype(np.float32)
    dyp_fold[..., 1, :] = RNG.integers(1, 9, (nseg, n_acc, n_freq, B))
    scheme = po.MiddleOut(nseg, 3, 10.0)
    level = 6
    idx_seg, coord_seg = scheme.get_segment_coords_so_far(level)
    coord_cur = scheme.get_coord(level)
    n = 120
    lv = np.zeros((n, P + 2, 2))
    lv[:, 0, 0] = RNG.uniform(-15, 15, n)
    lv[:, 1, 0] = RNG.uniform(-3e3, 3e3, n)
    lv[:, 2, 0] = RNG.uniform(-1e4, 1e4, n)
    lv[:, -1, 0] = RNG.uniform(199.5, 200.5, n)
    tree = po.Tree(lv, np.zeros((n, 2, B), np.float32), np.zeros(n, np.float32),
                   np.zeros((n, P + 2), np.int32))
    cfg = {"nbins": B, "counts_init": np.array([n_acc, n_freq]), "param_limits": limits,
           "widths": np.array([1, 2, 4, 8])}
    out = po.ascend(tree, dyp_fold, idx_seg, coord_seg, coord_cur, cfg)
    checked = 0
    for i in range(n):
        a, v, d0, f0 = (Fr(x) for x in lv[i, :, 0][[0, 1, 2, 3]])
        exp = np.zeros((2, B), dtype=np.int64)
        ok = True
        for iseg, (t0s, _) in zip(idx_seg, coord_seg):
            dt = Fr(t0s) - Fr(coord_cur[0])
            v_s = v + a * dt
            d_s = d0 + v * dt + a * dt * dt / 2
            freq = f0 * (1 - v_s / C)
            x = (dt - d_s / C) * f0
            pos = (x - math.floor(x)) * B
            ra = (a - Fr(-20)) * Fr(n_acc) / Fr(40) + Fr(1, 10**6)
            rf = (freq - Fr(199)) * Fr(n_freq) / Fr(2) + Fr(1, 10**6)
            if min(abs(ra - round(ra)), abs(rf - round(rf)),
                   abs(pos + Fr(1, 2) - round(pos + Fr(1, 2)))) < 1e-6:
                ok = False
                break
            ia = min(max(math.floor(ra), 0), n_acc - 1)
            jf = min(max(math.floor(rf), 0), n_freq - 1)
            s = math.floor(pos + Fr(1, 2)) % B
            exp += np.roll(dyp_fold[iseg, ia, jf].astype(np.int64), s, axis=1)
        if not ok:
            continue
        np.testing.assert_array_equal(out.folds[i], exp.astype(np.float32), err_msg=str(i))
        checked += 1
    assert checked > 100
    np.testing.assert_array_equal(out.scores, orc.snr_score_batch(out.folds, cfg["widths"]))
    np.testing.assert_array_equal(out.scores_ep, tree.scores)


def test_round_bf16_matches_torch_rne():
    """FP32 -> BF16 round to nearest even, against torch's conversion, over every sign,
    exponent and upper mantissa with the lower 16 bits at and around the tie."""
    torch = pytest.importorskip("torch")
    hi = np.arange(1 << 16, dtype=np.uint32)
    lows = np.array([0x0000, 0x0001, 0x7FFF, 0x8000, 0x8001, 0xFFFF], dtype=np.uint32)
    bits = (hi[:, None] << 16 | lows[None, :]).reshape(-1)
    x = bits.view(np.float32)
    keep = ~np.isnan(x)  # torch canonicalises NaN payloads
    x = x[keep]
    got = orc.round_bf16(x)
    exp = torch.from_numpy(x.copy()).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    np.testing.assert_array_equal(got, exp)
    np.testing.assert_array_equal(orc.bf16_to_f32(got), torch.from_numpy(
        x.copy()).to(torch.bfloat16).float().numpy())


def _decimal_snr(fold: np.ndarray, widths) -> np.ndarray:
    """Boxcar S/N (scoring.py:99-165 with e/sqrt(v)) of [n][2][B] integer rows, 50 digits."""
    getcontext().prec = 50
    n, _, B = fold.shape
    out = np.empty((n, len(widths)))
    for p in range(n):
        y = [Decimal(int(fold[p, 0, j])) / Decimal(int(fold[p, 1, j])).sqrt() for j in range(B)]
        ps, run = [], Decimal(0)
        for j in range(B):
            run += y[j]
            ps.append(run)
        ext = ps + [ps[j] + run for j in range(B)]
        for k, w in enumerate(int(x) for x in widths):
            h = (Decimal(B - w) / Decimal(B * w)).sqrt()
            b = Decimal(w) * h / Decimal(B - w)
            dmax = max(ext[i + w] - ext[i] for i in range(B))
            out[p, k] = float((h + b) * dmax - b * run)
    return out


def test_ffa_search_snr_f64_on_integer_data():
    """ffa_search_snr_f64: on small-integer samples every fold sum is exact in FP32 and
    FP64 alike, so its S/N must equal a high-precision evaluation of the FP32 oracle's
    (exact) final fold."""
    nsamps = 2**13
    lim = ParamLimits.from_taylor((195.0, 205.0), accel=(-50.0, 50.0)).limits
    cfg = PulsarSearchConfig(nsamps=nsamps, tsamp=64e-6, nbins=32, eta=1.0, param_limits=lim,
                             bseg_brute=1024, ducy_max=0.3, wtsp=1.5, use_fourier=False)
    ts_e = RNG.integers(-8, 9, nsamps).astype(np.float32)
    ts_v = RNG.integers(1, 4, nsamps).astype(np.float32)
    fold, pa = orc.ffa_fold(ts_e, ts_v, cfg)
    assert np.all(fold == np.round(fold))
    got = orc.ffa_search_snr_f64(ts_e, ts_v, cfg).reshape(-1, len(cfg.score_widths))
    sel = RNG.choice(fold.shape[1], min(200, fold.shape[1]), replace=False)
    exp = _decimal_snr(fold[0][sel], cfg.score_widths)
    np.testing.assert_allclose(got[sel], exp, rtol=1e-13, atol=1e-13)
    # The FP32 path on the same exact fold differs only by FP32 rounding.
    snr32 = orc.ffa_search_snr(ts_e, ts_v, cfg).reshape(got.shape)
    assert np.max(np.abs(snr32 - got)) < 1e-4 * np.max(np.abs(got))
