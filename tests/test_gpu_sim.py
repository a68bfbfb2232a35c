"""GPU parity of the seeded signal simulation (pyl_sim_series / pyl_sim_combine) against the
CPU oracle (oracle/sim_oracle.py).

The noise deviates and the fixed-point calibration folds are integer/bit work and must be
bit-exact; so must the template and the series for the derivative modulation, whose proper
time uses basic FP64 operations only. The orbital modulations call the device's sin / cos /
atan / tan (within an ulp of libm), so their templates are compared within 1e-6 and the
series within 1e-5 of its peak. Then a generated series is searched with ffa_search.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import sim_oracle as so
from pyloki_amd.simulation.pulse import PulseSignalConfig, build_cdf_lut

pytestmark = pytest.mark.gpu

CASES = [
    ("derivative", {"acc": 20.0, "jerk": 0.5}, 0.00431, 2**18 + 3, True),
    ("derivative", {}, 0.5, 2**16, True),  # 7812 bins: the fold without LDS staging
    ("derivative_series", {"coeffs": [0.0, 0.0, 20.0, 0.5, 1e-3, 1e-6]}, 0.00431, 2**17, False),
    ("circular", {"p_orb": 6 * 3600.0, "psi": 1.1, "x_orb": 1.5}, 0.004, 2**18, False),
    ("circular_t0", {"a": 4e8, "p_orb": 3 * 3600.0, "t0": 100.0}, 0.004, 2**17, False),
    ("keplerian", {"p_orb": 8 * 3600.0, "ecc": 0.3, "phi": 0.5, "a": 4e8, "om": 1.0,
                   "inc": 1.2}, 0.004, 2**17, False),
]


@pytest.mark.parametrize("kind,kw,period,n,exact", CASES)
def test_simulation_matches_the_oracle(gpu_device, kind, kw, period, n, exact):
    dt = 64e-6
    cfg = PulseSignalConfig(period=period, dt=dt, nsamps=n, snr=15.0, ducy=0.08,
                            mod_type=kind, mod_kwargs=kw)
    print(f"sim case {kind} n={n} period={period}", flush=True)
    ts_e, ts_v, info = cfg.simulate_device(seed=1234, stream_id=0, max_iter=8, tol=1e-3)
    lut = build_cdf_lut("gaussian", 0.08, 0.5, max(4096, int(100 / 0.08)))
    want_e, want_v, parts = so.generate(period, dt, n, 15.0, 0.08, kind, kw, cfg.mod_tref, lut,
                                        seed=1234, stream=0, max_iter=8, tol=1e-3)
    folded = info["folded"]
    if exact:
        assert np.array_equal(folded["var"], parts["folds"][2].astype(np.float32))
        assert np.array_equal(folded["noise"],
                              (parts["folds"][0] * 2.0**-32).astype(np.float32))
        assert np.array_equal(folded["template"],
                              (parts["folds"][1] * 2.0**-32).astype(np.float32))
        assert info["noise_scale"] == float(parts["noise_scale"])
        assert info["signal_scale"] == parts["signal_scale"]
        assert np.array_equal(ts_e.cpu().numpy(), want_e)
        assert np.array_equal(ts_v.cpu().numpy(), want_v)
    else:
        # A bin may move where a proper time sits within an ulp of a bin edge.
        moved = int(np.sum(np.abs(folded["var"] - parts["folds"][2])))
        assert moved <= 2
        got = ts_e.cpu().numpy()
        assert np.max(np.abs(got - want_e)) <= 1e-5 * np.max(np.abs(want_e))
        assert abs(info["signal_scale"] / parts["signal_scale"] - 1) < 1e-4


def test_simulation_noise_is_the_generators_row(gpu_device):
    """The noise equals pyl_rng_normal's row 0 of the same seed and stream (bit for bit),
    and a second call with the same seed reproduces every output bit."""
    import ctypes

    import torch

    from pyloki_amd import _lib
    from pyloki_amd.device import ptr, stream_ptr

    cfg = PulseSignalConfig(period=0.004, dt=64e-6, nsamps=2**16, snr=10.0, ducy=0.1)
    a = cfg.simulate_device(seed=77, stream_id=5)
    b = cfg.simulate_device(seed=77, stream_id=5)
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
    z = torch.empty(2**16, dtype=torch.float32, device=gpu_device)
    _lib.check(_lib.load().pyl_rng_normal(ctypes.c_uint64(77), ctypes.c_uint64(5), 1, 2**16, 1.0,
                                          ptr(z), stream_ptr()))
    tp = so.proper_time("derivative", {}, np.arange(2**16) * 64e-6, cfg.mod_tref)
    tmpl = so.template(tp, 64e-6, 0.004, build_cdf_lut("gaussian", 0.1, 0.5, 4096))
    want = (np.float32(a[2]["noise_scale"]) * z.cpu().numpy()).astype(np.float32) + (
        np.float32(a[2]["signal_scale"]) * tmpl).astype(np.float32)
    assert np.array_equal(a[0].cpu().numpy(), want)


def test_generated_series_is_found_by_the_search(gpu_device):
    """A series from the GPU generator (accelerated pulsar, S/N 20) is recovered by the
    FFA search at its frequency and acceleration."""
    from pyloki_amd.config import ParamLimits, PulsarSearchConfig
    from pyloki_amd.search import ffa_search

    n = 2**20
    f0 = 200.37
    ts = PulseSignalConfig(period=1 / f0, dt=64e-6, nsamps=n, snr=20.0, ducy=0.08,
                           mod_kwargs={"acc": 12.0}).generate(seed=3)
    lim = ParamLimits.from_taylor((195.0, 205.0), accel=(-50.0, 50.0)).limits
    cfg = PulsarSearchConfig(nsamps=n, tsamp=64e-6, nbins=64, eta=1.0, param_limits=lim,
                             bseg_brute=4096, use_fourier=False)
    _, pgram = ffa_search(ts, cfg)
    best = pgram.find_best_params()
    assert abs(best["freq"] - f0) < 0.02
    accels = np.asarray(pgram.params["accel"])
    step = float(np.max(np.diff(accels)))
    assert abs(best["accel"] - 12.0) <= step  # the nearest grid accelerations
    assert best["snr"] > 12.0
