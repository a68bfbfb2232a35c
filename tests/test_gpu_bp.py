"""Branching patterns on the device (pyl_bp_counts) against the host generators, which
the notebook goldens pin (tests/test_planning.py): equal bit for bit for the Taylor
kinds of the accel and jerk notebooks, every tiling strategy, moving and fixed grid, and
for the circular kinds of the circular-orbit notebook."""

from __future__ import annotations

import numpy as np
import pytest

from pyloki_amd.config import ParamLimits, PulsarSearchConfig
from tests.test_planning import GOLD, _circ_cfg

pytestmark = pytest.mark.gpu


def _taylor_cfg(name, tiling):
    g = GOLD[name]
    fu = g["from_upper"]
    p = ParamLimits.from_upper(tuple(fu["freq"]), fu["true_d_params"],
                               tuple(fu["d_upper_limits"]), g["nsamps"] * g["tsamp"])
    return PulsarSearchConfig(
        nsamps=g["nsamps"], tsamp=g["tsamp"], nbins=g["nbins"], eta=g["eta"],
        param_limits=p.limits, bseg_brute=g["bseg_brute"], bseg_ffa=g["bseg_ffa"],
        prune_poly_order=g["prune_poly_order"], ducy_max=g["ducy_max"], wtsp=g["wtsp"],
        branch_max=g["branch_max"], tiling_strategy=tiling)


@pytest.mark.parametrize("name,ref", [("ep_accel", 0), ("ep_jerk", 64)])
@pytest.mark.parametrize("tiling", ["aggressive", "conservative", "quadrature"])
@pytest.mark.parametrize("kind", ["poly_taylor_moving", "poly_taylor_fixed"])
def test_taylor_pattern_on_device(gpu_device, name, ref, tiling, kind):
    print(f"[case] {name} {kind} {tiling}", flush=True)
    cfg = _taylor_cfg(name, tiling)
    host = cfg.generate_branching_pattern(kind, ref)
    dev = cfg.generate_branching_pattern(kind, ref, device=True)
    np.testing.assert_array_equal(dev, host)


@pytest.mark.parametrize("kind", ["circ_taylor_moving", "circ_taylor_fixed"])
def test_circular_pattern_on_device(gpu_device, kind):
    print(f"[case] ep_circular {kind}", flush=True)
    g = GOLD["ep_circular"]
    cfg = _circ_cfg(g, np.array(g["param_limits_8dp"]))
    host = cfg.generate_branching_pattern(kind, 0)
    dev = cfg.generate_branching_pattern(kind, 0, device=True)
    np.testing.assert_array_equal(dev, host)
    if kind == "circ_taylor_moving":
        np.testing.assert_allclose(dev, g["branching_pattern_circ_taylor_moving_ref0"],
                                   rtol=6e-9, atol=5e-9)
    with pytest.raises(ValueError, match="Taylor and circular"):
        cfg.generate_branching_pattern("poly_chebyshev_moving", 0, device=True)
