"""The host-side Chebyshev pieces of the package (tables, transforms, seed, report,
branching pattern) agree bit for bit with the oracle, and the Chebyshev level constants of
the HIP library are those tables (CPU only: the library's host code is not run here)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import cheby_oracle as co
from pyloki_amd import prune
from pyloki_amd.config import ParamLimits, PulsarSearchConfig
from pyloki_amd.utils import maths, transforms

RNG = np.random.default_rng(5)


def test_tables_equal_the_oracle():
    for K in range(7):
        np.testing.assert_array_equal(maths.compute_connection_matrix_s(K), co.matrix_s(K))
        np.testing.assert_array_equal(maths.compute_connection_matrix_r(K), co.matrix_r(K))
    for _ in range(20):
        tc1, tc2 = RNG.uniform(-50, 50, 2)
        ts1, ts2 = RNG.uniform(0.1, 40, 2)
        for K in (2, 3, 4, 5):
            np.testing.assert_array_equal(
                maths.poly_chebyshev_transform_matrix(K, tc1, ts1, tc2, ts2, 1),
                co.transform_matrix(K, tc1, ts1, tc2, ts2))
    with pytest.raises(ValueError, match="must be positive"):
        maths.poly_chebyshev_transform_matrix(2, 0.0, 0.0, 1.0, 1.0)


@pytest.mark.parametrize("tiling", ["aggressive", "quadrature", "conservative"])
def test_transforms_equal_the_oracle(tiling):
    K = 3
    a = np.stack([RNG.normal(size=(30, K + 1)) * 1e3, RNG.uniform(1, 1e3, (30, K + 1))], -1)
    np.testing.assert_array_equal(transforms.taylor_to_cheby_full(a, 7.5),
                                  co.taylor_to_cheby_full(a, 7.5))
    np.testing.assert_array_equal(transforms.cheby_to_taylor_full(a, 7.5),
                                  co.cheby_to_taylor_full(a, 7.5))
    np.testing.assert_array_equal(transforms.shift_cheby_full(a, (12.0, 9.0), (10.0, 6.0),
                                                              tiling),
                                  co.shift_cheby_full(a, (12.0, 9.0), (10.0, 6.0), tiling))
    np.testing.assert_array_equal(transforms.cheby_to_taylor_param_shift(a[..., 0], 10.0, 6.0,
                                                                         3.25),
                                  co.cheby_param_shift(a[..., 0], 10.0, 6.0, 3.25))
    np.testing.assert_array_equal(transforms.taylor_to_cheby_errors(a[..., 1], 2.0),
                                  co.taylor_to_cheby_errors(a[..., 1], 2.0))


def test_seed_and_report_equal_the_oracle():
    pa = [np.linspace(-20, 20, 7), np.linspace(199.0, 201.0, 11)]
    dp = np.array([6.0, 0.003])
    for P, pa_p, dp_p in ((2, pa, dp), (3, [np.array([-1.0, 1.0]), *pa], np.r_[0.5, dp])):
        seed = prune.poly_chebyshev_seed(pa_p, dp_p, P, (40.0, 12.5))
        np.testing.assert_array_equal(seed, co.cheby_seed(pa_p, dp_p, P, (40.0, 12.5)))
        lv = seed.copy()
        lv[:, :-1, 0] += RNG.normal(size=lv[:, :-1, 0].shape) * 1e3
        np.testing.assert_array_equal(prune.poly_chebyshev_report_batch(lv, (50.0, 60.0)),
                                      co.cheby_report_batch(lv, (50.0, 60.0)))


@pytest.mark.parametrize("ref", [0, 3, 7])
def test_branching_pattern_equals_the_oracle(ref):
    nsamps = 2**20
    lim = ParamLimits.from_taylor((195.0, 205.0), accel=(-50.0, 50.0)).limits
    cfg = PulsarSearchConfig(nsamps=nsamps, tsamp=64e-6, nbins=32, eta=1.0, param_limits=lim,
                             bseg_brute=nsamps // 256, bseg_ffa=nsamps // 8,
                             prune_poly_order=2, ducy_max=0.3, wtsp=1.5)
    got = cfg.generate_branching_pattern("poly_chebyshev_moving", ref_seg=ref)
    dp = cfg.get_dparams(cfg.niters_ffa, True)
    exp = co.generate_bp_cheby(cfg.get_param_arr(dp), cfg.get_dparams_actual(cfg.niters_ffa),
                               cfg.tseg_ffa, 8, cfg.nbins, cfg.eta, ref, cfg.tiling_strategy)
    np.testing.assert_array_equal(got, exp)
    assert got.shape == (7,) and np.all(got >= 1)
    fixed = cfg.generate_branching_pattern("poly_chebyshev_fixed", ref_seg=ref)
    assert fixed.shape == (7,) and np.all(fixed >= 1)
    with pytest.raises(ValueError, match="Invalid kind"):
        cfg.generate_branching_pattern("poly_cheby_moving")
