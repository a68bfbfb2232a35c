"""The EP search captured as a HIP graph and replayed for several series of one
configuration (``graph.CapturedSearch``) against eager ``Pruning.execute`` on each series:
reports, scores and level statistics identical, replay after replay, with eager work and
synchronisations between the replays; also when a level is handed back to the host
(survivor overflow) and with a survivor budget inside the loop."""

from __future__ import annotations

import numpy as np
import pytest

from pyloki_amd.graph import CapturedSearch
from pyloki_amd.prune import Pruning
from tests.test_gpu_cheby import _setup
from tests.test_gpu_circ import _setup as _circ_setup

pytestmark = pytest.mark.gpu


def _case(msg: str) -> None:
    print(f"[case] {msg}", flush=True)


def _same(a: dict, b: dict, what: str) -> None:
    np.testing.assert_array_equal(a["param_sets"], b["param_sets"], err_msg=what)
    np.testing.assert_array_equal(a["scores"], b["scores"], err_msg=what)
    np.testing.assert_array_equal(a["scores_ep"], b["scores_ep"], err_msg=what)
    assert a["level_stats"] == b["level_stats"], what


def _replays(dyps, thresholds, ref, hook_factory=None, **kw):
    cs = CapturedSearch(dyps[0], thresholds, ref,
                        level_hook=hook_factory() if hook_factory else None, **kw)
    order = [0, 1, 2, 0, 1]
    for k, i in enumerate(order):
        pr = Pruning(dyps[i], thresholds, **kw)
        pr.level_hook = hook_factory() if hook_factory else None
        exp = pr.execute(ref)  # eager work between the replays
        got = cs.run(dyps[i])
        _same(got, exp, f"replay {k} (series {i})")
    return cs


@pytest.mark.parametrize("max_sugg", [2**20, 64])
def test_graph_taylor_replays(gpu_device, max_sugg):
    _case(f"graph taylor max_sugg={max_sugg}")
    dyps = [_setup(2, 2**17, 8, seed=s)[1] for s in (4, 5, 6)]
    thresholds = np.linspace(2.5, 5.5, 7)
    cs = _replays(dyps, thresholds, 3, max_sugg=max_sugg, batch_size=64)
    print(f"[result] captures {cs.captures}, hand-offs {cs.pr.loop_handoffs}", flush=True)
    if max_sugg == 64:
        assert cs.pr.loop_handoffs  # the overflow rule ran on the host


@pytest.mark.parametrize("kind", ["none", "snr", "cnn"])
def test_graph_circular_replays(gpu_device, kind):
    from pyloki_amd.cnn import BranchRanker, SurvivorBudget

    _case(f"graph circular complex budget={kind}")
    dyps = []
    for s in (3, 4, 5):
        cfg, dyp, _ = _circ_setup(use_fourier=True)
        if s != 3:  # other series of the same configuration
            from pyloki_amd.ffa import DynamicProgramming
            from pyloki_amd.simulation.pulse import PulseSignalConfig

            sim = PulseSignalConfig(period=1 / 200.17, dt=cfg.tsamp, nsamps=cfg.nsamps,
                                    snr=30.0, ducy=0.1,
                                    mod_type="circular",
                                    mod_kwargs={"p_orb": 30.0, "x_orb": 0.002, "psi": 0.1 * s})
            dyp = DynamicProgramming(sim.generate_analytic(seed=s), cfg)
            dyp.initialize()
            dyp.execute()
        dyps.append(dyp)
    thresholds = np.linspace(2.0, 8.0, dyps[0].nsegments - 1)
    ranker = BranchRanker(cfg.nbins, seed=2)
    factory = None
    if kind == "snr":
        factory = lambda: SurvivorBudget(512)  # noqa: E731
    elif kind == "cnn":
        factory = lambda: SurvivorBudget(512, ranker)  # noqa: E731
    cs = _replays(dyps, thresholds, dyps[0].nsegments // 2, factory, max_sugg=2**16,
                  batch_size=256)
    print(f"[result] captures {cs.captures}, hand-offs {cs.pr.loop_handoffs}", flush=True)


def test_search_batch_equals_eager(gpu_device):
    from pyloki_amd.cnn import SurvivorBudget
    from pyloki_amd.graph import SearchBatch

    _case("search batch (streams) vs eager")
    dyps = [_setup(2, 2**17, 8, seed=s)[1] for s in (4, 5, 6, 7, 8)]
    thresholds = np.linspace(2.5, 5.5, 7)
    for factory in (None, lambda: SurvivorBudget(24)):
        sb = SearchBatch(dyps[0], thresholds, 3, width=3, hook_factory=factory,
                         max_sugg=2**20, batch_size=64)
        got = sb.run(dyps)
        for i, d in enumerate(dyps):
            pr = Pruning(d, thresholds, max_sugg=2**20, batch_size=64)
            pr.level_hook = factory() if factory else None
            _same(got[i], pr.execute(3), f"series {i}")


def test_captured_ffa_batch_replays(gpu_device):
    """The batched FFA step (``graph.CapturedFFABatch``: brute fold, every level two at a
    time, the fused last levels and scoring) captured once and replayed: every replay's
    S/N grids equal an eager ``ffa_batch`` of the same series bit for bit, with eager work
    and synchronisations between the replays and new samples copied in between."""
    import torch

    import bench
    from pyloki_amd.ffa import FFAPlan, ffa_batch
    from pyloki_amd.graph import CapturedFFABatch

    _case("captured FFA batch (3 series of 2^18..2^20 samples)")
    lens = [2**20, 2**18, 2**19]
    cfg = bench.make_config(max(lens))
    plan = FFAPlan(cfg)
    widths = cfg.score_widths

    def series(seed):
        r = np.random.default_rng(seed)
        return [(torch.from_numpy(r.standard_normal(n).astype(np.float32)).cuda(),
                 torch.ones(n, dtype=torch.float32, device="cuda")) for n in lens]

    base = series(1)
    cap = CapturedFFABatch(plan, base, widths)
    for k, seed in enumerate([1, 1, 2, 3, 2]):
        data = series(seed)
        exp = [x.cpu().numpy() for x in ffa_batch(plan, data, widths)]  # eager between
        torch.cuda.synchronize()
        got = [x.cpu().numpy() for x in cap.run(data)]
        for i in range(len(lens)):
            np.testing.assert_array_equal(got[i], exp[i], err_msg=f"replay {k} series {i}")
    print("[result] 5 replays bit-equal to eager ffa_batch", flush=True)
