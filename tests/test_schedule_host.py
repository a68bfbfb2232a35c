"""Host logic of the survivor-budget schedules: ``budget_schedule`` and the leave-one-out
recipe choice of ``select_schedule`` (no GPU)."""

from __future__ import annotations

import numpy as np

from pyloki_amd import cnn

# Three calibration series over levels 1..2; series C needs 5000 at level 2 (an outlier).
NEEDS = [{1: 10, 2: 100}, {1: 20, 2: 50}, {1: 15, 2: 5000}]
SURV = [{1: 30, 2: 10000}] * 3


def test_budget_schedule_max_margin_floor_window():
    s = cnn.budget_schedule(NEEDS, 2, margin=1.5, floor=1, window=0)
    np.testing.assert_array_equal(s, [1, 30, 7500])
    s = cnn.budget_schedule(NEEDS, 2, margin=1.0, floor=40, window=1)
    np.testing.assert_array_equal(s, [40, 5000, 5000])


def test_select_schedule_prefers_leave_one_out_coverage():
    # margin 1.0: leaving B out gives 15 < 20 at level 1, leaving C out 100 < 5000: 1 kept;
    # margin 1.5: B is kept too (ceil(1.5 * 15) = 23 >= 20): 2 kept.
    r = cnn.select_schedule(NEEDS, SURV, 2, margins=(1.0, 1.5), floors=(1,), windows=(0,))
    assert (r["margin"], r["loo_kept"], r["loo_of"]) == (1.5, 2, 3)
    np.testing.assert_array_equal(r["schedule"], [1, 30, 7500])


def test_select_schedule_breaks_ties_by_kept_survivors():
    # margins 1.5 and 2.0 both keep 2 of 3; 1.5 keeps 3 x (30 + 7500) survivors, 2.0 keeps
    # 3 x (30 + 10000).
    r = cnn.select_schedule(NEEDS, SURV, 2, margins=(2.0, 1.5), floors=(1,), windows=(0,))
    assert r["margin"] == 1.5
    assert r["kept_survivors"] == 3 * (30 + 7500)
