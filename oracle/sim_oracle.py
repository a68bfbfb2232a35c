"""CPU oracle of the seeded signal simulation -- TEST INFRASTRUCTURE ONLY.

Restates ``simulation/pulse.py:310-406`` (PulseSignalConfig.generate), the template
integral ``generate_pulse_template`` (:39-66), the calibration ``calibrate_scale_on_folds``
(:69-111) and the proper time of ``simulation/modulate.py`` (:34-54, :92-97, :147-150,
:222-226, :240-250 with kepler.py:64-140), vectorised in NumPy FP64, on the generator of
the C oracle (``thresh_oracle.rng_normal``: the deviates the product draws for row 0 of a
stream). The calibration folds are the product's documented fixed-point sums (value
* 2^32 rounded to nearest even, summed in int64), formed here with exact integer
arithmetic (``np.add.at``).

Pinning (tests/test_sim_oracle.py): every template value against scipy's exact CDF
difference within the table's interpolation error; the sum over a turn against the
closed form (the CDF telescopes); the proper time of each modulation against independent
closed forms (a derivative series of the same coefficients, the Taylor expansion of the
orbit, a circular orbit for a Keplerian one with zero eccentricity); the fixed-point folds
against FP64 bin sums; the calibration by its defining property (the folded S/N of the
series at the true ephemeris is the target within ``tol``); the deviates through the
Philox known-answer vectors of the threshold oracle's tests.
"""

from __future__ import annotations

import numpy as np

from oracle import thresh_oracle

C_VAL = 299792458.0


def _py_mod(a, b):
    """Python's float modulo (divisor's sign), elementwise."""
    return np.mod(a, b)


def proper_time(kind: str, kw: dict, t: np.ndarray, t_ref: float) -> np.ndarray:
    if kind == "derivative":
        s, v, a, j, sn = (kw.get(k, 0.0) for k in ("shift", "vel", "acc", "jerk", "snap"))
        x = t - t_ref
        delay = s + x * (v + x * (a * 0.5 + x * (j * (1.0 / 6.0) + x * sn * (1.0 / 24.0))))
        return t - delay / C_VAL
    if kind == "derivative_series":
        c = np.asarray(kw["coeffs"], dtype=np.float64)
        fact = np.array([float(np.prod(np.arange(1, k + 1))) for k in range(len(c))])
        x = t - t_ref
        delay = np.zeros_like(t)
        for k in range(len(c)):
            delay = delay + (c[k] / fact[k]) * x ** k
        return t - delay / C_VAL
    if kind == "circular":
        omega = 2 * np.pi / kw["p_orb"]
        return t - kw["x_orb"] * np.sin(omega * (t - t_ref) + kw["psi"])
    if kind == "circular_t0":
        omega = 2 * np.pi / kw["p_orb"]
        phi = 2 * np.pi * _py_mod(t_ref - kw["t0"], kw["p_orb"]) / kw["p_orb"]
        return t - kw["a"] * np.sin(omega * (t - t_ref) + phi) / C_VAL
    if kind == "keplerian":
        p, ecc = kw["p_orb"], abs(kw["ecc"])
        m = 2 * np.pi * (_py_mod(t - t_ref, p) / p) + kw["phi"]
        e = m.copy()
        for _ in range(8):
            e += (m + ecc * np.sin(e) - e) / (1 - ecc * np.cos(e))
        nu = 2 * np.arctan(np.sqrt((1 + ecc) / (1 - ecc)) * np.tan(e / 2))
        r = kw["a"] * (1 - ecc**2) / (1 + ecc * np.cos(nu))
        return t - r * np.sin(nu + kw["om"]) * np.sin(kw["inc"]) / C_VAL
    raise ValueError(kind)


def template(tp: np.ndarray, dt: float, period: float, lut: np.ndarray) -> np.ndarray:
    """pulse.py:39-66: FP64 CDF-table interpolation at the phases of t' and t' + dt."""
    ngrid = len(lut) - 1
    lut64 = lut.astype(np.float64)

    def cdf(p):
        x = p * ngrid
        i = x.astype(np.int64)
        f = x - i
        top = i >= ngrid
        i = np.minimum(i, ngrid - 1)
        val = (1.0 - f) * lut64[i] + f * lut64[i + 1]
        return np.where(top, lut64[ngrid], val)

    inv = 1.0 / period
    p0 = _py_mod(tp, period) * inv
    p1 = _py_mod(tp + dt, period) * inv
    out = np.where(p1 > p0, cdf(p1) - cdf(p0), 0.0)
    return out.astype(np.float32)


def fold_bins(tp: np.ndarray, freq: float, nbins: int) -> np.ndarray:
    """get_phase_idx_int(t', freq, nbins, 0) elementwise (psr_utils.py:12-56)."""
    ph = tp * freq
    ph = ph - np.floor(ph)
    ip = ph * nbins
    ip[ip >= nbins] = 0.0
    b = (ip.astype(np.float32) + np.float32(0.5)).astype(np.int64)
    b[b == nbins] = 0
    return b


def fixed_fold(x: np.ndarray, bins: np.ndarray, nbins: int) -> np.ndarray:
    acc = np.zeros(nbins, dtype=np.int64)
    np.add.at(acc, bins, np.rint(x.astype(np.float64) * 2.0**32).astype(np.int64))
    return acc


def boxcar_snr_1d(x: np.ndarray, widths: np.ndarray) -> np.ndarray:
    """scoring.py:105-140 (stdnoise 1) for widths below the profile length."""
    x = np.asarray(x, dtype=np.float32)
    n = len(x)
    ps = np.empty(n + int(np.max(widths)), dtype=np.float32)
    run = np.float32(0.0)
    for i in range(n):
        run = np.float32(run + x[i])
        ps[i] = run
    for j in range(len(ps) - n):
        ps[n + j] = np.float32(ps[j] + run)
    out = np.empty(len(widths), dtype=np.float32)
    for iw, w in enumerate(widths):
        w = int(w)
        h = np.sqrt((n - w) / (n * w))
        b = w * h / (n - w)
        dmax = np.max(ps[w:w + n] - ps[:n])
        out[iw] = (h + b) * float(dmax) - b * float(run)
    return out


def calibrate(noise_vals, noise_var, tmpl_vals, snr, initial, max_iter, tol, damping=0.8):
    """pulse.py:69-111."""
    nbins = len(noise_vals)
    widths = np.arange(1, nbins // 2)
    tn = (tmpl_vals / np.sqrt(noise_var)).astype(np.float32)
    nn = (noise_vals / np.sqrt(noise_var)).astype(np.float32)
    s_t = float(np.max(boxcar_snr_1d(tn, widths)))
    scale = float(initial)
    for _ in range(max_iter):
        comb = np.array([float(nn[i]) + scale * float(tn[i]) for i in range(nbins)],
                        dtype=np.float64).astype(np.float32)
        diff = float(np.max(boxcar_snr_1d(comb, widths))) - snr
        if diff * diff < tol * tol:
            break
        scale -= damping * diff / s_t
    return scale


def generate(period, dt, nsamps, snr, ducy, kind, kw, t_ref, lut, seed, stream=0,
             phi0=0.5, max_iter=5, tol=1e-2):
    """pulse.py:310-406 with the product's seeded deviates and fixed-point folds. Returns
    (ts_e, ts_v, parts) where parts holds the template, noise, folds and scales."""
    t = np.arange(nsamps, dtype=np.float64) * dt
    tp = proper_time(kind, kw, t, t_ref)
    tmpl = template(tp, dt, period, lut)
    nq = (nsamps + 3) // 4
    noise = thresh_oracle.rng_normal(seed, stream, 1, 4 * nq)[0, :nsamps]
    nbins = int(period / dt)
    bins = fold_bins(tp, 1.0 / period, nbins)
    acc_n = fixed_fold(noise, bins, nbins)
    acc_t = fixed_fold(tmpl, bins, nbins)
    cnt = np.bincount(bins, minlength=nbins).astype(np.int64)
    nv = (acc_n.astype(np.float64) * 2.0**-32).astype(np.float32)
    tv = (acc_t.astype(np.float64) * 2.0**-32).astype(np.float32)
    var = cnt.astype(np.float32)
    center = round(phi0 * nbins) % nbins
    width = max(1, round(ducy * nbins))
    off = np.ones(nbins, dtype=bool)
    off[(center - width // 2 + np.arange(width)) % nbins] = False
    ns = np.float32(1.0) / np.std(nv[off])
    var_s = (var * ns**2).astype(np.float32)
    tn = (tv / np.sqrt(var_s)).astype(np.float32)
    s_t = np.max(boxcar_snr_1d(tn, np.arange(1, nbins // 2)))
    initial = (np.float32(snr) / s_t) * ns
    cs = calibrate((nv * ns).astype(np.float32), var_s, tv, snr, float(initial), max_iter, tol)
    ts_e = (ns * noise).astype(np.float32) + (np.float32(cs) * tmpl).astype(np.float32)
    ts_v = np.full(nsamps, ns**2, dtype=np.float32)
    parts = {"proper_time": tp, "template": tmpl, "noise": noise, "bins": bins,
             "folds": np.stack([acc_n, acc_t, cnt]), "noise_scale": ns, "signal_scale": cs}
    return ts_e, ts_v, parts
