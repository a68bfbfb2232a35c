"""GPU parity of the complex (use_fourier=True) FFA against the oracle, within FP32
tolerance: the leaf fold (a GEMM whose summation order is the library's), every level
(complex64 phasor recurrences), the irfft scorer, and the whole search's S/N grid
(<= 1e-5 of the peak S/N)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import complex_oracle as co
from oracle import oracle as orc
from pyloki_amd import kernels
from pyloki_amd.config import ParamLimits, PulsarSearchConfig
from pyloki_amd.ffa import DynamicProgramming, FFAPlan
from pyloki_amd.search import ffa_search
from pyloki_amd.simulation.pulse import PulseSignalConfig

pytestmark = pytest.mark.gpu


def _case(msg: str) -> None:
    print(f"[case] {msg}", flush=True)


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _cfg(P: int, nsamps: int, nbins: int) -> PulsarSearchConfig:
    if P == 1:
        lim = np.array([[1.8, 2.2]])
        tsamp = 256e-6
    else:
        lim = ParamLimits.from_taylor((195.0, 205.0), accel=(-50.0, 50.0)).limits
        tsamp = 64e-6
    return PulsarSearchConfig(nsamps=nsamps, tsamp=tsamp, nbins=nbins, eta=1.0,
                              param_limits=lim, ducy_max=0.3, wtsp=1.5, use_fourier=True)


def _series(cfg: PulsarSearchConfig, seed: int = 2):
    period = 0.5 if cfg.nparams == 1 else 1 / 200.37
    mod = {} if cfg.nparams == 1 else {"acc": 12.0}
    return PulseSignalConfig(period=period, dt=cfg.tsamp, nsamps=cfg.nsamps, snr=25.0,
                             ducy=0.08, mod_kwargs=mod).generate_analytic(seed=seed)


@pytest.mark.parametrize("nbins,seg_len,nfreq,nseg", [
    (64, 4096, 37, 6), (32, 2048, 9, 6), (128, 1024, 5, 6),
    # several segment and comb tiles with tails; a K tail (1000 = 62.5 stages of 16); rows
    # that are not 16-byte aligned (1002 samples: the scalar-load path)
    (64, 1000, 11, 70), (32, 1002, 3, 9), (64, 2048, 40, 130)])
def test_complex_leaf_fold(gpu_device, nbins, seg_len, nfreq, nseg):
    _case(f"complex leaf fold nbins={nbins} seg_len={seg_len} nfreq={nfreq} nseg={nseg}")
    rng = np.random.default_rng(nbins + nseg)
    ts_e = rng.normal(size=nseg * seg_len).astype(np.float32)
    ts_v = rng.uniform(0.5, 1.5, nseg * seg_len).astype(np.float32)
    freqs = np.linspace(150.0, 210.0, nfreq)
    t_ref = seg_len * 64e-6 / 2
    plan = kernels.ComplexFoldPlan(freqs, seg_len, nbins, 64e-6, t_ref)
    got = plan.fold(_dev(ts_e), _dev(ts_v)).cpu().numpy()
    exp = co.brutefold_complex(ts_e, ts_v, freqs, seg_len, nbins, 64e-6, t_ref)
    scale_e = np.abs(ts_e.reshape(nseg, seg_len)).sum(axis=1)[:, None, None]
    scale_v = np.abs(ts_v.reshape(nseg, seg_len)).sum(axis=1)[:, None, None]
    err_e = np.abs(got[:, :, 0] - exp[:, :, 0]) / scale_e
    err_v = np.abs(got[:, :, 1] - exp[:, :, 1]) / scale_v
    print(f"[result] leaf fold max error / sum|x|: e {err_e.max():.2e}, v {err_v.max():.2e}",
          flush=True)
    assert err_e.max() < 2e-6 and err_v.max() < 2e-6


@pytest.mark.parametrize("P,nsamps,nbins", [(1, 2**18, 64), (2, 2**18, 64), (2, 2**17, 32)])
def test_complex_levels(gpu_device, P, nsamps, nbins):
    _case(f"complex levels P={P} nsamps={nsamps} nbins={nbins}")
    cfg = _cfg(P, nsamps, nbins)
    ts = _series(cfg)
    plan = FFAPlan(cfg)
    got = plan.run(_dev(ts.ts_e), _dev(ts.ts_v)).cpu().numpy()
    exp, _ = co.ffa_fold_complex(ts.ts_e, ts.ts_v, cfg)
    assert got.shape == exp.shape
    scale = np.abs(exp).max(axis=-1, keepdims=True)
    rel = np.abs(got - exp).max() / np.abs(exp).max()
    print(f"[result] complex levels max error / max |X|: {rel:.2e}", flush=True)
    assert np.all(np.abs(got - exp) <= 1e-5 * np.maximum(scale, 1e-3 * np.abs(exp).max()))


def test_complex_scorer(gpu_device):
    _case("complex scorer")
    rng = np.random.default_rng(4)
    nbins, n = 64, 500
    prof = rng.normal(size=(n, 2, nbins)).astype(np.float32)
    prof[:, 1] = rng.uniform(50.0, 60.0, (n, nbins))
    folds = np.fft.rfft(prof).astype(np.complex64)
    widths = np.array([1, 2, 3, 5, 8, 13])
    got = kernels.score_boxcar_complex(_dev(folds), widths, nbins).cpu().numpy()
    exp = orc.boxcar_snr(co.fold_norm_complex(folds, nbins), widths)
    err = np.abs(got - exp).max() / np.abs(exp).max()
    print(f"[result] complex scorer max error / max S/N: {err:.2e}", flush=True)
    assert err <= 1e-5


@pytest.mark.parametrize("P", [1, 2])
def test_complex_search_snr(gpu_device, P):
    _case(f"complex ffa_search P={P}")
    cfg = _cfg(P, 2**18, 64)
    ts = _series(cfg)
    dyp, pgram = ffa_search(ts, cfg)
    exp = co.ffa_search_snr_complex(ts.ts_e, ts.ts_v, cfg)
    got = pgram.snrs.reshape(exp.shape)
    peak = float(np.abs(exp).max())
    err = float(np.abs(got - exp).max()) / peak
    print(f"[result] complex search S/N max error / peak: {err:.2e} (peak {peak:.2f})",
          flush=True)
    assert err <= 1e-5
    assert np.unravel_index(np.argmax(got), got.shape) == np.unravel_index(np.argmax(exp),
                                                                           exp.shape)
    # DynamicProgramming keeps the reference's complex layout and normalisation
    assert dyp.fold.dtype == np.complex64 and dyp.fold.shape[-1] == cfg.nbins // 2 + 1
    norm = dyp.get_fold_norm()
    assert norm.shape[-1] == cfg.nbins



# ------------------------------------------------------------------ EP with complex folds
# Each level's oracle input is the GPU's previous tree, capped at _EP_CAP parents (the same
# capped tree goes to both), so every level is checked at a bounded CPU cost. Leaves and
# backtracks come from the FP64 path and must agree (bit-exactly for Taylor cells, within
# 1e-12 of the column's scale for cells moved along the circular orbit); scores and folds
# are FP32 complex arithmetic and must agree within 1e-5 of the level's largest value; a
# survivor may be present on one side only if its score is within 1e-5 of the threshold.

_EP_CAP = 4096
_TOL = 1e-5


def _ep_setup(circular: bool, use_fourier: bool = True):
    if circular:
        nsamps, nseg, p_orb_min = 2**18, 32, 20.0
        lim = ParamLimits.from_circular((199.0, 201.0), p_orb_min=p_orb_min, m_c_max=0.2,
                                        m_p_min=1.4).limits
        extra = dict(prune_poly_order=5, p_orb_min=p_orb_min, m_c_max=0.2, m_p_min=1.4)
        mod = {"p_orb": 30.0, "x_orb": 0.002, "psi": 0.7}
        mod_type = "circular"
        period = 1 / 200.17
    else:
        nsamps, nseg = 2**18, 16
        lim = ParamLimits.from_taylor((195.0, 205.0), accel=(-50.0, 50.0)).limits
        extra = dict(prune_poly_order=2)
        mod = {"acc": 12.0}
        mod_type = "derivative"
        period = 1 / 200.37
    cfg = PulsarSearchConfig(nsamps=nsamps, tsamp=64e-6, nbins=32, eta=1.0, param_limits=lim,
                             bseg_brute=nsamps // 256, bseg_ffa=nsamps // nseg, ducy_max=0.3,
                             wtsp=1.5, use_fourier=use_fourier, branch_max=32, **extra)
    ts = PulseSignalConfig(period=period, dt=64e-6, nsamps=nsamps, snr=30.0, ducy=0.1,
                           mod_type=mod_type,
                           mod_kwargs=mod).generate_analytic(seed=3 if circular else 4)
    dyp = DynamicProgramming(ts, cfg)
    dyp.initialize()
    dyp.execute()
    pcfg = {"poly_order": cfg.prune_poly_order, "nbins": cfg.nbins, "eta": cfg.eta,
            "branch_max": cfg.branch_max,
            "counts_init": np.array([len(a) for a in dyp.param_arr]),
            "param_limits": cfg.param_limits, "widths": cfg.score_widths,
            "tiling": cfg.tiling_strategy, "prop_sig": cfg.propagator_significance,
            "val_sig": cfg.validation_significance, "p_orb_min": cfg.p_orb_min,
            "x_mass_const": cfg.x_mass_const}
    thresholds = (np.linspace(2.0, 8.0, nseg - 1) if circular
                  else np.linspace(2.0, 5.0, nseg - 1))
    return cfg, dyp, pcfg, thresholds


def _host_tree(t):
    from oracle import prune_oracle as po

    return po.Tree(t.leaves.cpu().numpy(), t.folds.cpu().numpy(), t.scores.cpu().numpy(),
                   t.backtracks.cpu().numpy())


def _cap(pr) -> None:
    from pyloki_amd.prune import DeviceTree

    t = pr.world_tree
    if t.size > _EP_CAP:
        pr._world_tree = DeviceTree(t.leaves[:_EP_CAP].contiguous(),
                                    t.folds[:_EP_CAP].contiguous(),
                                    t.scores[:_EP_CAP].contiguous(),
                                    t.backtracks[:_EP_CAP].contiguous())


def _same_row(g, e, i, j, scale) -> bool:
    if not np.array_equal(g.backtracks[i], e.backtracks[j]):
        return False
    if not np.array_equal(g.leaves[i, -1], e.leaves[j, -1]):
        return False
    return bool(np.all(np.abs(g.leaves[i, :-1] - e.leaves[j, :-1]) <= 1e-12 * scale[0]))


def _align(g, e, thr: float, what: str):
    """Row pairs (i, j) of the GPU tree g and the oracle tree e, both in child order;
    a row on one side only must score within _TOL of the threshold."""
    scale = (np.maximum(np.abs(e.leaves[:, :-1]).max(axis=0, keepdims=True), 1e-300)
             if e.size else None)
    near = _TOL * max(abs(thr), 1.0)
    pairs, i, j, skipped = [], 0, 0, 0
    while i < g.size or j < e.size:
        if i < g.size and j < e.size and _same_row(g, e, i, j, scale):
            pairs.append((i, j))
            i += 1
            j += 1
        elif j < e.size and abs(float(e.scores[j]) - thr) <= near:
            j += 1
            skipped += 1
        elif i < g.size and abs(float(g.scores[i]) - thr) <= near:
            i += 1
            skipped += 1
        else:
            raise AssertionError(f"{what}: survivors differ at GPU row {i}, oracle row {j}")
    return np.array(pairs, dtype=np.int64).reshape(-1, 2), skipped


def _close(got, exp, what: str) -> float:
    if exp.size == 0:
        return 0.0
    err = float(np.max(np.abs(got - exp))) / max(float(np.max(np.abs(exp))), 1e-30)
    assert err <= _TOL, f"{what}: {err:.2e}"
    return err


@pytest.mark.parametrize("circular", [False, True])
def test_complex_prune_levels(gpu_device, circular):
    from oracle import prune_oracle as po
    from pyloki_amd.prune import Pruning

    _case(f"complex EP circular={circular}")
    cfg, dyp, pcfg, thresholds = _ep_setup(circular)
    nseg = dyp.nsegments
    ref = nseg // 2 if circular else 3
    max_sugg = 2**20
    pr = Pruning(dyp, thresholds, max_sugg=max_sugg, batch_size=256)
    pr.initialize(ref)
    scheme = po.MiddleOut(nseg, ref, dyp.tseg)
    fold = dyp.fold
    nb_f = cfg.nbins // 2 + 1
    seed = np.ascontiguousarray(fold[scheme.ref_idx]).reshape(-1, 2, nb_f)
    _close(pr.world_tree.scores.cpu().numpy(), co.cscore_batch(seed, cfg.score_widths), "seed")
    level_fn = po.circ_prune_level if circular else po.prune_level
    worst, skipped, flags, n_phy_cut = 0.0, 0, set(), False
    for level in range(1, nseg):
        _cap(pr)
        tin = _host_tree(pr.world_tree)
        pr.execute_iter()
        coords = {"init": scheme.get_coord(0), "cur": scheme.get_current_coord(level),
                  "next": scheme.get_coord(level), "add": scheme.get_segment_coord(level)}
        thr = float(thresholds[level - 1])
        exp, st, _ = level_fn(tin, fold[scheme.get_segment_idx(level)], coords, thr, pcfg,
                              max_sugg, 256)
        assert exp.size < max_sugg
        got = _host_tree(pr.world_tree)
        pairs, sk = _align(got, exp, thr, f"level {level}")
        skipped += sk
        gi, ej = pairs[:, 0], pairs[:, 1]
        worst = max(worst, _close(got.scores[gi], exp.scores[ej], f"level {level} scores"))
        _close(got.folds[gi], exp.folds[ej], f"level {level} folds")
        gst = pr.pstats[level]
        assert (gst.n_leaves, gst.n_leaves_phy) == (st.n_leaves, st.n_leaves_phy), level
        flags |= set(np.unique(exp.leaves[:, -1, 1]).tolist())
        n_phy_cut |= st.n_leaves_phy < st.n_leaves
        if pr.prune_complete:
            break
    print(f"[result] complex EP levels {pr.prune_level}: worst score error / max "
          f"{worst:.2e}, near-threshold survivors on one side {skipped}, flags "
          f"{sorted(flags)}", flush=True)
    assert pr.prune_level > nseg // 2
    if circular:
        assert 1.0 in flags and n_phy_cut
    _cap(pr)
    tin = _host_tree(pr.world_tree)
    pr._do_ascend()
    lvl = pr.prune_level
    idx_seg, coord_seg = scheme.get_segment_coords_so_far(lvl)
    asc = po.circ_ascend if circular else po.ascend
    exp = asc(tin, fold, idx_seg, coord_seg, scheme.get_coord(lvl), pcfg)
    _close(pr.world_tree.folds.cpu().numpy(), exp.folds, "ascend folds")
    _close(pr.world_tree.scores.cpu().numpy(), exp.scores, "ascend scores")
    np.testing.assert_array_equal(pr.world_tree.scores_ep.cpu().numpy(), tin.scores)


def test_complex_prune_execute(gpu_device):
    """The whole complex-fold EP run finds the real-fold run's candidate."""
    from pyloki_amd.prune import Pruning

    _case("complex EP execute + report")
    best = {}
    for cplx in (False, True):
        cfg, dyp, _, thresholds = _ep_setup(False, use_fourier=cplx)
        res = Pruning(dyp, thresholds, max_sugg=2**18).execute(3)
        b = int(np.argmax(res["scores"]))
        best[cplx] = (res["param_sets"][b], float(res["scores"][b]))
    (rep_r, s_r), (rep_c, s_c) = best[False], best[True]
    print(f"[result] best S/N real {s_r:.3f} complex {s_c:.3f}; f {rep_r[-3, 0]:.5f} "
          f"{rep_c[-3, 0]:.5f}; a {rep_r[0, 0]:.3f} {rep_c[0, 0]:.3f}", flush=True)
    assert abs(rep_c[-3, 0] - 200.37) < 0.05
    assert abs(rep_c[-3, 0] - rep_r[-3, 0]) < 0.01
    assert abs(rep_c[0, 0] - rep_r[0, 0]) < 2.0
    assert abs(s_c - s_r) / s_r < 0.15
