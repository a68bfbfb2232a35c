"""Pin the CPU oracle of the seeded signal simulation (oracle/sim_oracle.py) and the host
logic of pyloki_amd.simulation (no GPU): template values against scipy's exact CDF, the
telescoping sum over a turn, proper times against independent closed forms, the
fixed-point folds against FP64 sums, the calibration by its defining property, and the
host calibration of the product against the oracle's restatement."""

from __future__ import annotations

import numpy as np
import pytest
from scipy import stats

from oracle import sim_oracle as so
from pyloki_amd.simulation import modulate, pulse

C = 299792458.0


def _lut(shape="gaussian", width=0.1, pos=0.5, ngrid=4096):
    return pulse.build_cdf_lut(shape, width, pos, ngrid)


@pytest.mark.parametrize("shape", ["gaussian", "boxcar", "von_mises"])
def test_template_matches_the_exact_cdf_difference(shape):
    """Each value is F(p1) - F(p0) of the shape's distribution up to the table's linear
    interpolation error (and 0 across the turn boundary)."""
    period, dt, width = 0.00431, 64e-6, 0.1
    n = 4096
    t = np.arange(n) * dt
    lut = _lut(shape, width)
    got = so.template(t, dt, period, lut).astype(np.float64)
    p0 = np.mod(t, period) / period
    p1 = np.mod(t + dt, period) / period
    if shape == "gaussian":
        rv = stats.norm(loc=0.5, scale=width / (2 * np.sqrt(2 * np.log(10))))
    elif shape == "boxcar":
        rv = stats.uniform(loc=0.5, scale=width)
    else:
        rv = stats.vonmises(loc=0.5, kappa=np.log(2.0) / (2.0 * np.sin(np.pi * width / 2.0) ** 2))
    exact = np.where(p1 > p0, rv.cdf(p1) - rv.cdf(p0), 0.0)
    # The last table cell interpolates towards the wrap value F(0); elsewhere the linear
    # interpolation error is below 1e-5 for these widths at 4096 cells.
    inner = (p1 < 1 - 1.0 / 4096) & (p0 < 1 - 1.0 / 4096)
    tol = 2e-3 if shape == "boxcar" else 1e-5  # the boxcar CDF has kinks inside cells
    assert np.max(np.abs(got - exact)[inner]) < tol
    assert np.all(got[p1 <= p0] == 0)


def test_template_sum_over_a_turn_telescopes():
    """Inside one turn each value is c(p_{i+1}) - c(p_i): the sum is c(last) - c(first)."""
    period, dt = 0.0125, 50e-6  # 250 samples a turn
    t = np.arange(1000) * dt
    lut = _lut("gaussian", 0.08)
    tmpl = so.template(t, dt, period, lut).astype(np.float64)
    p = np.mod(t, period) / period
    start = np.nonzero(np.diff(p) < 0)[0][0] + 1  # first sample of the second turn
    stop = start + 249  # the turn's last full step ends at phase 249/250
    s = tmpl[start:stop].sum()
    ngrid = len(lut) - 1

    def c(q):
        x = q * ngrid
        i = int(x)
        f = x - i
        return (1 - f) * float(lut[i]) + f * float(lut[i + 1])

    assert abs(s - (c(p[stop]) - c(p[start]))) < 1e-6
    assert abs(s - 1.0) < 1e-6  # the pulse at phase 0.5 lies inside the turn


def test_proper_time_closed_forms():
    t = np.linspace(0, 500.0, 2001)
    tref = 250.0
    # Derivative model vs the same coefficients as a derivative series.
    kw = {"shift": 3.0, "vel": 1e3, "acc": 20.0, "jerk": 0.5, "snap": 1e-3}
    a = so.proper_time("derivative", kw, t, tref)
    b = so.proper_time("derivative_series", {"coeffs": [3.0, 1e3, 20.0, 0.5, 1e-3]}, t, tref)
    assert np.max(np.abs(a - b)) < 1e-12
    x = t - tref
    exact = t - (3.0 + 1e3 * x + 10.0 * x**2 + 0.5 * x**3 / 6 + 1e-3 * x**4 / 24) / C
    assert np.max(np.abs(a - exact)) < 1e-12
    # Circular orbit vs its Taylor expansion about t_ref (small omega x).
    circ = {"p_orb": 86400.0, "psi": 0.3, "x_orb": 1.2}
    cm = modulate.CircularModulating(**circ)
    d = cm.to_derivatives()
    coeffs = [d["shift"], d["vel"], d["acc"], d["jerk"], d["snap"], d["crackle"]]
    ser = so.proper_time("derivative_series", {"coeffs": coeffs}, t, tref)
    assert np.max(np.abs(so.proper_time("circular", circ, t, tref) - ser)) < 1e-11
    # Keplerian with zero eccentricity is a circular orbit: z = a sin(M + om) sin(inc).
    kep = {"p_orb": 3600.0, "ecc": 0.0, "phi": 0.4, "a": 3e8, "om": 0.2, "inc": np.pi / 2}
    m = 2 * np.pi * np.mod(x, 3600.0) / 3600.0 + 0.4
    want = t - 3e8 * np.sin(m + 0.2) / C
    assert np.max(np.abs(so.proper_time("keplerian", kep, t, tref) - want)) < 1e-12
    # Circular with an epoch t0 = t_ref - psi / omega is the circular orbit with phase psi.
    p_orb = 7200.0
    t0 = tref - 0.3 / (2 * np.pi / p_orb)
    ct0 = so.proper_time("circular_t0", {"a": 1.2 * C, "p_orb": p_orb, "t0": t0}, t, tref)
    cc = so.proper_time("circular", {"p_orb": p_orb, "psi": 0.3, "x_orb": 1.2}, t, tref)
    assert np.max(np.abs(ct0 - cc)) < 1e-9


@pytest.mark.parametrize("kind,kw", [
    ("derivative", {"acc": 20.0, "jerk": 0.5}),
    ("derivative_series", {"coeffs": [0.0, 0.0, 20.0, 0.5, 1e-3, 1e-6]}),
    ("circular", {"p_orb": 6 * 3600.0, "psi": 1.1, "x_orb": 1.5}),
    ("circular_t0", {"a": 4e8, "p_orb": 3 * 3600.0, "t0": 100.0}),
    ("keplerian", {"p_orb": 8 * 3600.0, "ecc": 0.3, "phi": 0.5, "a": 4e8, "om": 1.0,
                   "inc": 1.2}),
])
def test_host_modulations_equal_the_oracle(kind, kw):
    t = np.arange(4096) * 64e-6 * 97
    tref = t[-1] / 2
    got = modulate.type_to_mods[kind](**kw).generate(t, tref)
    want = so.proper_time(kind, kw, t, tref)
    assert np.max(np.abs(got - want)) <= 2e-12


def test_fixed_point_folds_match_fp64_sums():
    rng = np.random.default_rng(3)
    x = rng.standard_normal(100_000).astype(np.float32)
    bins = rng.integers(0, 61, x.size)
    acc = so.fixed_fold(x, bins, 61)
    ref = np.bincount(bins, weights=x.astype(np.float64), minlength=61)
    counts = np.bincount(bins, minlength=61)
    assert np.all(np.abs(acc * 2.0**-32 - ref) <= counts * 2.0**-33 + 1e-12)


def test_fold_bins_match_the_phase_index():
    from oracle import oracle as orc

    rng = np.random.default_rng(5)
    tp = rng.uniform(-5, 500, 2000)
    b = so.fold_bins(tp, 1 / 0.00437, 68)
    assert all(int(b[i]) == orc.phase_idx_int(float(tp[i]), 1 / 0.00437, 68, 0.0)
               for i in range(len(tp)))


def test_boxcar_matches_the_oracle_scorer():
    from oracle import oracle as orc

    rng = np.random.default_rng(7)
    x = rng.standard_normal(61).astype(np.float32)
    w = np.arange(1, 30)
    assert np.array_equal(so.boxcar_snr_1d(x, w), orc.boxcar_snr(x, w).astype(np.float32))
    assert np.array_equal(pulse.boxcar_snr_1d(x, w), so.boxcar_snr_1d(x, w))


@pytest.mark.parametrize("kind,kw,snr", [
    ("derivative", {"acc": 15.0}, 12.0),
    ("circular", {"p_orb": 1800.0, "psi": 0.4, "x_orb": 0.01}, 25.0),
])
def test_oracle_generate_reaches_the_target_snr(kind, kw, snr):
    """The defining property of generate(): the series folded at the true ephemeris
    (FP64 bin sums of ts_e and ts_v) has the target best boxcar S/N within tol."""
    period, dt, n = 0.00411, 64e-6, 2**16
    lut = _lut("gaussian", 0.1)
    ts_e, ts_v, parts = so.generate(period, dt, n, snr, 0.1, kind, kw, n * dt / 2, lut,
                                    seed=11, max_iter=10, tol=1e-2)
    nb = int(period / dt)
    b = parts["bins"]
    fe = np.bincount(b, weights=ts_e.astype(np.float64), minlength=nb)
    fv = np.bincount(b, weights=ts_v.astype(np.float64), minlength=nb)
    prof = (fe / np.sqrt(fv)).astype(np.float32)
    got = float(np.max(so.boxcar_snr_1d(prof, np.arange(1, nb // 2))))
    assert abs(got - snr) < 0.05
    assert ts_v[0] == np.float32(parts["noise_scale"]) ** 2


def test_host_calibration_equals_the_oracle():
    """The product's host calibration (pulse.calibration_scales) gives the oracle's noise
    and signal scales on the same folded profiles."""
    rng = np.random.default_rng(9)
    nb = 64
    nv = rng.standard_normal(nb).astype(np.float32) * 30
    var = np.full(nb, 1000.0, dtype=np.float32)
    tv = (pulse.generate_folded_profile(nb, 0.1) * 40).astype(np.float32)
    ns, cs, _ = pulse.calibration_scales(nv, var, tv, 20.0, 0.5, 0.1, 5, 1e-2)
    off = pulse.off_pulse_mask(nb, 0.5, 0.1)
    ns_o = np.float32(1.0) / np.std(nv[off])
    var_s = (var * ns_o**2).astype(np.float32)
    tn = (tv / np.sqrt(var_s)).astype(np.float32)
    s_t = np.max(so.boxcar_snr_1d(tn, np.arange(1, nb // 2)))
    cs_o = so.calibrate((nv * ns_o).astype(np.float32), var_s, tv, 20.0,
                        float((np.float32(20.0) / s_t) * ns_o), 5, 1e-2)
    assert ns == ns_o
    assert cs == cs_o


def test_config_keeps_the_reference_fields():
    cfg = pulse.PulseSignalConfig(period=0.004, dt=64e-6, nsamps=2**12, snr=10, ducy=0.05,
                                  mod_type="circular",
                                  mod_kwargs={"p_orb": 3600.0, "psi": 0.1, "x_orb": 0.5})
    assert cfg.mod_tref == cfg.tobs / 2
    assert cfg.fold_bins_ideal == int(0.004 / 64e-6)
    assert cfg.get_updated({"snr": 20, "bogus": 1}).snr == 20
    with pytest.raises(ValueError, match="Duty cycle"):
        pulse.PulseSignalConfig(period=0.004, dt=64e-6, ducy=1.0)
    with pytest.raises(TypeError):
        pulse.PulseSignalConfig(period=0.004, dt=64e-6, mod_kwargs={"p_orb": 1.0})
    m = cfg.mod_func.to_device(cfg.mod_tref)
    assert m.kind == modulate.PYL_MOD_CIRCULAR and m.p[0] == 3600.0 and m.p[2] == 0.5
