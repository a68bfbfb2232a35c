"""GPU parity of the fixed-grid EP search (use_moving_grid=False: fixed coordinates,
fixed resolve, identity transform) against the oracle, for the Taylor, Chebyshev and
circular bases: every level's tree bit-exactly (circular leaf values within 1e-12, as for
the moving grid), the ascend, the report, and the sync-free level loop against the
per-level path."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import cheby_oracle as co
from oracle import prune_oracle as po
from pyloki_amd.prune import Pruning
from tests.test_gpu_circ import _compare as _circ_compare
from tests.test_gpu_circ import _setup as _circ_setup
from tests.test_gpu_eploop import _run, _same
from tests.test_gpu_prune import _compare_tree, _setup

pytestmark = pytest.mark.gpu


def _case(msg: str) -> None:
    print(f"[case] {msg}", flush=True)


def _levels(pr, exp, nseg, compare) -> None:
    for level in range(1, nseg):
        pr.execute_iter()
        ot = exp["trees"][level - 1] if level - 1 < len(exp["trees"]) else None
        if ot is None:
            assert pr.prune_complete
            break
        compare(pr.world_tree, ot, f"level {level}")
        st, ost = pr.pstats[level], exp["stats"][level - 1]
        assert st.n_leaves == ost.n_leaves
        assert np.float32(st.score_max) == np.float32(ost.score_max)
        if pr.prune_complete:
            break
    pr._do_ascend()
    t = exp["tree"]
    if t.size:
        np.testing.assert_array_equal(pr.world_tree.folds.cpu().numpy(), t.folds)
        np.testing.assert_array_equal(pr.world_tree.scores.cpu().numpy(), t.scores)
        np.testing.assert_array_equal(pr.world_tree.scores_ep.cpu().numpy(), t.scores_ep)


@pytest.mark.parametrize("basis,P,nseg,thr,max_sugg,ref", [
    ("taylor", 2, 8, 2.5, 2**20, 3),
    ("taylor", 2, 16, 2.0, 300, 7),     # survivor buffer overflows
    ("taylor", 3, 8, 3.0, 2**20, 0),
    ("chebyshev", 2, 8, 2.5, 2**20, 3),
    ("chebyshev", 3, 8, 3.0, 2**20, 5),
])
def test_fixed_grid_levels_bit_exact(gpu_device, basis, P, nseg, thr, max_sugg, ref):
    _case(f"fixed grid {basis} P={P} nseg={nseg} max_sugg={max_sugg} ref={ref}")
    cfg, dyp, pcfg = _setup(P, 2**17, nseg)
    thresholds = np.linspace(thr, thr + 3.0, nseg - 1)
    run = po.prune_run if basis == "taylor" else co.cheby_prune_run
    exp = run(dyp.fold, dyp.param_arr, dyp.dparams_actual, dyp.tseg, thresholds, ref, pcfg,
              max_sugg, batch_size=64, moving=False)
    pr = Pruning(dyp, thresholds, max_sugg=max_sugg, batch_size=64, poly_basis=basis,
                 use_moving_grid=False)
    pr.initialize(ref)
    _levels(pr, exp, nseg, _compare_tree)
    res = Pruning(dyp, thresholds, max_sugg=max_sugg, batch_size=64, poly_basis=basis,
                  use_moving_grid=False).execute(ref)
    np.testing.assert_array_equal(res["param_sets"], exp["report"])
    np.testing.assert_array_equal(res["scores"], exp["tree"].scores)
    if max_sugg == 2**20 and P == 2:
        best = int(np.argmax(res["scores"]))
        assert abs(res["param_sets"][best, -3, 0] - 200.37) < 0.05


@pytest.mark.parametrize("basis,fourier,max_sugg", [
    ("taylor", False, 2**20), ("taylor", False, 64), ("taylor", True, 2**20),
    ("chebyshev", True, 2**20)])
def test_fixed_grid_loop_equals_per_level_path(gpu_device, basis, fourier, max_sugg):
    _case(f"fixed grid loop {basis} fourier={fourier} max_sugg={max_sugg}")
    cfg, dyp, pcfg = _setup(2, 2**18, 16) if not fourier else _fourier_setup()
    thresholds = np.linspace(2.0, 5.0, dyp.nsegments - 1)
    kw = {"max_sugg": max_sugg, "batch_size": 64, "poly_basis": basis,
          "use_moving_grid": False}
    _same(_run(dyp, thresholds, 7, True, **kw), _run(dyp, thresholds, 7, False, **kw))


def _fourier_setup():
    from tests.test_gpu_cheby import _setup as cheby_setup

    return cheby_setup(2, 2**18, 16, use_fourier=True)


def test_fixed_grid_circular_levels(gpu_device):
    _case("fixed grid circular, 2^17 samples, 16 segments")
    cfg, dyp, pcfg = _circ_setup(nsamps=2**17, nseg=16)
    nseg = dyp.nsegments
    thresholds = np.linspace(2.0, 7.0, nseg - 1)
    ref = nseg // 2
    exp = po.circ_prune_run(dyp.fold, dyp.param_arr, dyp.dparams_actual, dyp.tseg, thresholds,
                            ref, pcfg, 2**16, batch_size=1024, moving=False)
    pr = Pruning(dyp, thresholds, max_sugg=2**16, batch_size=1024, use_moving_grid=False)
    pr.initialize(ref)

    def compare(g, o, what):
        _circ_compare(g, o, what)

    _levels(pr, exp, nseg, compare)
    fast = _run(dyp, thresholds, ref, True, max_sugg=2**16, use_moving_grid=False)
    slow = _run(dyp, thresholds, ref, False, max_sugg=2**16, use_moving_grid=False)
    _same(fast, slow)
    np.testing.assert_allclose(fast[1]["param_sets"], exp["report"], rtol=1e-12, atol=1e-9)
    assert len(fast[1]["scores"]) > 0
