"""GPU parity of the FFA path (brute fold, resolve, shift-add, scoring) against the oracle.

Integer/index results and FP32 folds are compared bit-exactly; S/N is compared bit-exactly
as well (the tolerance the path promises is 1e-5 relative; the kernels reproduce the
oracle's operation order, so equality is expected and asserted).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from pyloki_amd import _lib, kernels
from pyloki_amd.config import ParamLimits, PulsarSearchConfig
from pyloki_amd.ffa import DynamicProgramming, FFAPlan, ffa_batch
from pyloki_amd.io.timeseries import TimeSeries
from pyloki_amd.search import ffa_search
from pyloki_amd.simulation.pulse import PulseSignalConfig

pytestmark = pytest.mark.gpu


def _dev(x: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _case(msg: str) -> None:
    print(f"[case] {msg}", flush=True)


@pytest.mark.parametrize("nbins,seg_len,nfreq,t_ref_half,nsamps", [
    (64, 8192, 37, False, 8192 * 6),          # LDS-staged segment, t_ref = 0
    (64, 4096, 84, True, 4096 * 4 + 1000),    # partial last segment, t_ref = T/2
    (32, 16384, 12, False, 16384 * 3),        # segment too large for LDS (global reads)
    (100, 2048, 9, True, 2048 * 5),           # nbins not a power of two
    (128, 1024, 20, False, 1024 * 8),
])
def test_brutefold_bit_exact(gpu_device, nbins, seg_len, nfreq, t_ref_half, nsamps):
    _case(f"brutefold nbins={nbins} seg_len={seg_len} nfreq={nfreq} nsamps={nsamps}")
    rng = np.random.default_rng(nbins + seg_len)
    ts_e = rng.normal(size=nsamps).astype(np.float32)
    ts_v = rng.uniform(0.5, 1.5, size=nsamps).astype(np.float32)
    tsamp = 64e-6 if seg_len != 16384 else 256e-6
    freqs = np.sort(rng.uniform(150.0, 210.0, size=nfreq)) if tsamp < 1e-4 else \
        np.linspace(1.5, 2.5, nfreq)
    t_ref = seg_len * tsamp / 2 if t_ref_half else 0.0
    plan = kernels.BruteFoldPlan(freqs, seg_len, nbins, tsamp, t_ref)
    got = plan.fold(_dev(ts_e), _dev(ts_v)).cpu().numpy()
    exp = orc.brutefold(ts_e, ts_v, freqs, seg_len, nbins, tsamp, t_ref)
    np.testing.assert_array_equal(got, exp)


def _cfg(P: int, nsamps: int, nbins: int, bseg: int) -> PulsarSearchConfig:
    if P == 1:
        lim = ParamLimits.from_taylor((195.0, 205.0)).limits
    elif P == 2:
        lim = ParamLimits.from_taylor((195.0, 205.0), accel=(-50.0, 50.0)).limits
    else:
        lim = ParamLimits.from_taylor((495.0, 505.0), accel=(-30.0, 30.0), jerk=(-2.0, 2.0)).limits
    return PulsarSearchConfig(nsamps=nsamps, tsamp=64e-6, nbins=nbins, eta=1.0, param_limits=lim,
                              bseg_brute=bseg, use_fourier=False)


@pytest.mark.parametrize("P", [1, 2, 3])
def test_resolve_table_bit_exact(gpu_device, P):
    _case(f"resolve P={P}")
    cfg = _cfg(P, 2**20, 64, 2048)
    counts_prev = cfg.get_param_grid_count(0)
    for level in range(1, cfg.niters_ffa + 1):
        pa = cfg.get_param_arr(cfg.get_dparams(level))
        counts = np.array([len(a) for a in pa])
        params = _dev(np.concatenate(pa))
        phases = torch.empty((int(np.prod(counts)), 2), dtype=torch.float64, device="cuda")
        table = kernels.ffa_resolve(params, counts, counts_prev, cfg.param_limits, level,
                                    cfg.tseg_brute, cfg.nbins, phases=phases).cpu().numpy()
        dummy = np.zeros((2, int(np.prod(counts_prev)), 2, cfg.nbins), dtype=np.float32)
        _, rows, ph = orc.unify_fold(dummy, counts_prev, pa, cfg.param_limits, level,
                                     cfg.tseg_brute, cfg.nbins, return_table=True)
        np.testing.assert_array_equal(table[:, 0], rows[:, 0])
        np.testing.assert_array_equal(table[:, 2], rows[:, 1])
        np.testing.assert_array_equal(phases.cpu().numpy(), ph)
        sh = np.vectorize(lambda x: orc.shift_int(x, cfg.nbins))(ph)
        np.testing.assert_array_equal(table[:, 1], sh[:, 0])
        np.testing.assert_array_equal(table[:, 3], sh[:, 1])
        counts_prev = counts


@pytest.mark.parametrize("P,nbins,nsamps,bseg", [
    (1, 64, 2**18, 4096),
    (2, 64, 2**19, 4096),
    (3, 64, 2**19, 4096),
    (2, 32, 2**18, 2048),
    (2, 48, 2**17, 2048),   # generic (non power-of-two) shift-add kernel
    (1, 128, 2**18, 4096),
])
def test_ffa_all_levels_bit_exact(gpu_device, P, nbins, nsamps, bseg):
    _case(f"ffa levels P={P} nbins={nbins} nsamps={nsamps}")
    cfg = _cfg(P, nsamps, nbins, bseg)
    rng = np.random.default_rng(P * 1000 + nbins)
    ts_e = rng.normal(size=nsamps).astype(np.float32)
    ts_v = np.ones(nsamps, dtype=np.float32)
    plan = FFAPlan(cfg)
    folds = plan.run(_dev(ts_e), _dev(ts_v), keep=True)
    exp, _ = orc.ffa_fold(ts_e, ts_v, cfg, levels=0)
    np.testing.assert_array_equal(folds[0].cpu().numpy(), exp)
    counts_prev = cfg.get_param_grid_count(0)
    for level in range(1, cfg.niters_ffa + 1):
        pa = cfg.get_param_arr(cfg.get_dparams(level))
        exp = orc.unify_fold(folds[level - 1].cpu().numpy(), counts_prev, pa, cfg.param_limits,
                             level, cfg.tseg_brute, cfg.nbins)
        np.testing.assert_array_equal(folds[level].cpu().numpy(), exp)
        counts_prev = cfg.get_param_grid_count(level)


@pytest.mark.parametrize("nbins,n", [(64, 1000), (64, 64), (32, 7), (100, 130), (128, 300)])
def test_score_bit_exact(gpu_device, nbins, n):
    _case(f"score nbins={nbins} n={n}")
    rng = np.random.default_rng(n)
    e = rng.normal(size=(n, nbins)).astype(np.float32) * 5
    v = rng.uniform(0.5, 50.0, size=(n, nbins)).astype(np.float32)
    ev = np.ascontiguousarray(np.stack([e, v], axis=1))
    widths = np.array([1, 2, 3, 5, 8, 13, nbins // 2 - 1], dtype=np.int64)
    got = kernels.score_boxcar(_dev(ev), widths).cpu().numpy()
    exp = orc.boxcar_snr(orc.fold_norm(ev), widths)
    np.testing.assert_array_equal(got, exp)
    got_max = kernels.score_boxcar(_dev(ev), widths, mode=_lib.PYL_SCORE_MAX).cpu().numpy()
    np.testing.assert_array_equal(got_max, orc.snr_score_batch(ev, widths))
    norm = orc.fold_norm(ev)
    got_n = kernels.score_boxcar(_dev(norm), widths, 1.0, layout=_lib.PYL_LAYOUT_NORM)
    np.testing.assert_array_equal(got_n.cpu().numpy(), exp)
    np.testing.assert_allclose(got, exp, rtol=1e-5)


def _width_set(nbins: int, kind: str) -> np.ndarray:
    if kind == "default":
        return np.array([1, 2, 3, 5, 8, nbins // 4], dtype=np.int64)
    if kind == "half":       # widest width in (16, nbins/2]
        return np.array([1, 4, nbins // 2], dtype=np.int64)
    if kind == "wide":       # widest width nbins - 1
        return np.array([1, 7, nbins // 2 + 3, nbins - 1], dtype=np.int64)
    return np.arange(1, 21, dtype=np.int64)  # "many": more than 16 widths


@pytest.mark.parametrize("nbins,kind", [(32, "default"), (64, "default"), (128, "default"),
                                        (32, "wide"), (64, "half"), (64, "wide"),
                                        (64, "many")])
def test_fused_last_level_score_bit_exact(gpu_device, nbins, kind):
    _case(f"fused level+score nbins={nbins} widths={kind}")
    cfg = _cfg(2, 2**18, nbins, 4096)
    rng = np.random.default_rng(nbins)
    ts_e = rng.normal(size=cfg.nsamps).astype(np.float32)
    plan = FFAPlan(cfg)
    folds = plan.run(_dev(ts_e), _dev(np.ones(cfg.nsamps, dtype=np.float32)), keep=True)
    lp = plan.levels[-1]
    widths = _width_set(nbins, kind)
    exp = orc.boxcar_snr(orc.fold_norm(folds[-1].cpu().numpy()), widths)
    for n_inner in (None, int(lp.counts_cur[-1])):   # 1-D rows and (outer x inner) tiles
        got = kernels.ffa_shift_add_score(folds[-2], lp.table, lp.ncoords, widths,
                                          n_inner=n_inner).cpu().numpy()
        np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("nbins", [32, 64, 128])
def test_wrap_layout_levels_bit_exact(gpu_device, nbins):
    """The wrap row layout ([.., 2, B + 4], each component followed by its first 4 bins) on
    either side of a level gives the dense level's values, and the fused last level reads
    it to the same S/N (rounds kernel and, for nbins 128 or many widths, the LDS kernel)."""
    _case(f"wrap layout nbins={nbins}")
    cfg = _cfg(2, 2**18, nbins, 4096)
    rng = np.random.default_rng(100 + nbins)
    ts_e = rng.normal(size=cfg.nsamps).astype(np.float32)
    plan = FFAPlan(cfg)
    folds = plan.run(_dev(ts_e), _dev(np.ones(cfg.nsamps, dtype=np.float32)), keep=True)
    lp_prev, lp = plan.levels[-2], plan.levels[-1]
    src = folds[-3]
    dense = kernels.ffa_shift_add(src, lp_prev.table, lp_prev.ncoords)
    wrap = kernels.ffa_shift_add(src, lp_prev.table, lp_prev.ncoords, out_layout="wrap")
    assert wrap.shape[-1] == nbins + 4
    np.testing.assert_array_equal(wrap[..., :nbins].cpu().numpy(), dense.cpu().numpy())
    np.testing.assert_array_equal(wrap[..., nbins:].cpu().numpy(), dense[..., :4].cpu().numpy())
    src_wrap = torch.cat([src, src[..., :4]], dim=-1).contiguous()
    for lay_out in ("dense", "wrap"):
        got = kernels.ffa_shift_add(src_wrap, lp_prev.table, lp_prev.ncoords, in_layout="wrap",
                                    out_layout=lay_out)
        np.testing.assert_array_equal(got[..., :nbins].cpu().numpy(), dense.cpu().numpy())
    for kind in ("default", "many"):
        widths = _width_set(nbins, kind)
        exp = kernels.ffa_shift_add_score(dense, lp.table, lp.ncoords, widths,
                                          n_inner=int(lp.counts_cur[-1])).cpu().numpy()
        np.testing.assert_array_equal(
            exp, orc.boxcar_snr(orc.fold_norm(folds[-1].cpu().numpy()), widths))
        got = kernels.ffa_shift_add_score(wrap, lp.table, lp.ncoords, widths,
                                          n_inner=int(lp.counts_cur[-1]),
                                          in_layout="wrap").cpu().numpy()
        np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("nbins", [32, 64])
def test_two_fused_levels_bit_exact(gpu_device, nbins):
    """The last two levels fused with scoring, from the fold two levels before the last
    (dense or wrap rows), equal the two levels run one after the other and the oracle."""
    _case(f"two fused levels nbins={nbins}")
    cfg = _cfg(2, 2**18, nbins, 4096)
    rng = np.random.default_rng(200 + nbins)
    ts_e = rng.normal(size=cfg.nsamps).astype(np.float32)
    ts_v = rng.uniform(0.5, 1.5, cfg.nsamps).astype(np.float32)
    plan = FFAPlan(cfg)
    folds = plan.run(_dev(ts_e), _dev(ts_v), keep=True)
    lp_prev, lp = plan.levels[-2], plan.levels[-1]
    src = folds[-3]
    assert src.shape[0] == 4
    src_wrap = torch.cat([src, src[..., :4]], dim=-1).contiguous()
    for kind in ("default", "half", "wide"):
        widths = _width_set(nbins, kind)
        exp = kernels.ffa_shift_add_score(folds[-2], lp.table, lp.ncoords, widths,
                                          n_inner=int(lp.counts_cur[-1])).cpu().numpy()
        if kind == "default":
            np.testing.assert_array_equal(
                exp, orc.boxcar_snr(orc.fold_norm(folds[-1].cpu().numpy()), widths))
        for part, lay in ((src, "dense"), (src_wrap, "wrap")):
            # packed (default) and unpacked gather descriptors
            for packed in (1, 0):
                _lib.check(_lib.load().pyl_ffa_set_packed_tables(packed))
                try:
                    got = kernels.ffa_shift_add2_score(part, lp_prev.table, lp.table,
                                                       lp_prev.ncoords, lp.ncoords, widths,
                                                       n_inner=int(lp.counts_cur[-1]),
                                                       in_layout=lay)
                finally:
                    _lib.check(_lib.load().pyl_ffa_set_packed_tables(1))
                np.testing.assert_array_equal(got.cpu().numpy(), exp)
    with pytest.raises(ValueError, match="at most 16 widths"):
        kernels.ffa_shift_add2_score(src, lp_prev.table, lp.table, lp_prev.ncoords, lp.ncoords,
                                     _width_set(nbins, "many"))


def test_score_empty_and_invalid(gpu_device):
    _case("score empty / invalid")
    empty = torch.empty((0, 2, 64), dtype=torch.float32, device="cuda")
    assert kernels.score_boxcar(empty, np.array([1, 2])).shape == (0, 2)
    with pytest.raises(ValueError, match="width"):
        kernels.score_boxcar(torch.ones((3, 2, 64), device="cuda"), np.array([64]))


def test_ffa_search_config1_full_size(gpu_device):
    """Plain FFA frequency search at its full size: 2^21 samples, tsamp 256 us, nbins 64,
    f in [1.5, 2.5] Hz, widths 1..32 -- folds and the S/N grid equal the oracle."""
    _case("ffa_search config 1 full size")
    nsamps, tsamp = 2**21, 256e-6
    sim = PulseSignalConfig(period=0.5, dt=tsamp, nsamps=nsamps, snr=15.0, ducy=0.05)
    ts = sim.generate_analytic(seed=11)
    cfg = PulsarSearchConfig(nsamps=nsamps, tsamp=tsamp, nbins=64, eta=1.0,
                             param_limits=np.array([[1.5, 2.5]]), ducy_max=0.5, wtsp=2.0,
                             use_fourier=False)
    dyp, pgram = ffa_search(ts, cfg)
    exp = orc.ffa_search_snr(ts.ts_e, ts.ts_v, cfg)
    np.testing.assert_array_equal(pgram.snrs, exp)
    best = pgram.find_best_params()
    assert abs(best["freq"] - 2.0) < 1e-3 and best["snr"] > 10
    fold_exp, _ = orc.ffa_fold(ts.ts_e, ts.ts_v, cfg)
    np.testing.assert_array_equal(dyp.fold.reshape(fold_exp.shape), fold_exp)


def test_dynamic_programming_api(gpu_device):
    _case("DynamicProgramming API")
    nsamps = 2**18
    ts = TimeSeries(np.random.default_rng(5).normal(size=nsamps), np.ones(nsamps), 64e-6)
    cfg = _cfg(2, nsamps, 64, 4096)
    dyp = DynamicProgramming(ts, cfg)
    dyp.initialize()
    assert dyp.fold.shape == (nsamps // 4096, 1, len(dyp.param_arr[-1]), 2, 64)
    dyp.execute()
    counts = [len(p) for p in dyp.param_arr]
    assert dyp.fold.shape == (1, *counts, 2, 64)
    assert dyp.ffa_level == cfg.niters_ffa
    np.testing.assert_allclose(dyp.tseg, cfg.tseg_ffa)
    norm = dyp.get_fold_norm()
    assert norm.shape == (*counts, 64)


@pytest.mark.parametrize("lens", [[2**17, 2**19, 2**18, 2**17],
                                  [2**14, 2**13, 2**12, 2**15, 2**16]])
def test_ffa_batch_matches_single_series(gpu_device, lens):
    """Mixed lengths, including series with 8, 4, 2 and 1 segments after the brute fold
    (the two-level, one-level and level-0 scoring paths)."""
    _case(f"ffa_batch mixed lengths {lens}")
    cfg = _cfg(2, 2**19, 64, 4096)
    plan = FFAPlan(cfg)
    rng = np.random.default_rng(9)
    series = []
    for n in lens:
        e = rng.normal(size=n).astype(np.float32)
        series.append((_dev(e), _dev(np.ones(n, dtype=np.float32))))
    widths = cfg.score_widths
    out = ffa_batch(plan, series, widths)
    out_unfused = ffa_batch(plan, series, widths, fused=False)
    for a, b in zip(out, out_unfused):
        np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
    for (e, v), snr, n in zip(series, out, lens):
        if n == 4096:  # one segment: scored at level 0 (no FFA config exists for it)
            continue
        sub = PulsarSearchConfig(nsamps=n, tsamp=cfg.tsamp, nbins=64, eta=1.0,
                                 param_limits=cfg.param_limits, bseg_brute=4096,
                                 use_fourier=False)
        exp = orc.ffa_search_snr(e.cpu().numpy(), v.cpu().numpy(), sub)
        np.testing.assert_array_equal(snr.cpu().numpy(), exp.reshape(-1, len(widths)))


@pytest.mark.parametrize("seg_len,nbins,segs", [
    (4096, 64, 0), (4096, 64, 1), (2048, 64, 4), (8192, 32, 0), (1000, 100, 2),
    (16384, 32, 0),   # beyond the LDS budget: the global path
    (70000, 64, 0),   # more than 65536 samples per segment: u32 sample lists
])
def test_brutefold_batch_bit_exact(gpu_device, seg_len, nbins, segs):
    """A batch of series of different lengths (partial last segments included) in one
    launch: every series' rows equal the oracle's fold of that series."""
    _case(f"brutefold batch seg_len={seg_len} nbins={nbins} segs_per_block={segs}")
    lib = _lib.load()
    rng = np.random.default_rng(seg_len + nbins)
    lens = [seg_len * 5 + 123, seg_len * 3, seg_len * 7 - 5, seg_len, 17] if seg_len < 70000 \
        else [seg_len * 2 + 9, seg_len]
    tsamp = 64e-6
    freqs = np.sort(rng.uniform(150.0, 210.0, size=11))
    t_ref = seg_len * tsamp / 2
    plan = kernels.BruteFoldPlan(freqs, seg_len, nbins, tsamp, t_ref)
    host = [(rng.normal(size=n).astype(np.float32), rng.uniform(0.5, 1.5, n).astype(np.float32))
            for n in lens]
    dev = [(_dev(e), _dev(v)) for e, v in host]
    nseg = [-(-n // seg_len) for n in lens]
    _lib.check(lib.pyl_brutefold_set_segments_per_block(segs))
    try:
        out = torch.empty((sum(nseg), len(freqs), 2, nbins), dtype=torch.float32, device="cuda")
        plan.fold_batch(dev, out)
        out16 = torch.empty_like(out, dtype=torch.bfloat16)
        plan.fold_batch(dev, out16)
    finally:
        _lib.check(lib.pyl_brutefold_set_segments_per_block(0))
    got = out.cpu().numpy()
    got16 = out16.view(torch.int16).cpu().numpy().astype(np.uint16)
    off = 0
    for (e, v), n in zip(host, nseg):
        exp = orc.brutefold(e, v, freqs, seg_len, nbins, tsamp, t_ref)
        np.testing.assert_array_equal(got[off:off + n], exp)
        np.testing.assert_array_equal(got16[off:off + n], orc.round_bf16(exp))
        off += n
    assert plan.boundary_count >= 0


@pytest.mark.parametrize("P,nbins", [(2, 64), (2, 32), (3, 128), (1, 256)])
def test_shift_add2_equals_two_levels(gpu_device, P, nbins):
    """Two levels in one pass (dense and wrap output) equal two single levels bit for bit,
    at every pair of consecutive levels of a small search."""
    _case(f"shift_add2 P={P} nbins={nbins}")
    nsamps = 2**17
    cfg = _cfg(P, nsamps, nbins, nsamps // 64)
    plan = FFAPlan(cfg)
    rng = np.random.default_rng(P * nbins)
    e = _dev(rng.normal(size=nsamps).astype(np.float32))
    v = _dev(rng.uniform(0.5, 1.5, nsamps).astype(np.float32))
    folds = plan.run(e, v, keep=True)
    for lv in range(0, len(plan.levels) - 1):
        if folds[lv].shape[0] % 4:
            continue
        a, b = plan.levels[lv], plan.levels[lv + 1]
        got = kernels.ffa_shift_add2(folds[lv], a.table, b.table, a.ncoords, b.ncoords)
        np.testing.assert_array_equal(got.cpu().numpy(), folds[lv + 2].cpu().numpy())
        gw = kernels.ffa_shift_add2(folds[lv], a.table, b.table, a.ncoords, b.ncoords,
                                    out_layout="wrap").cpu().numpy()
        exp = folds[lv + 2].cpu().numpy()
        np.testing.assert_array_equal(gw[..., :nbins], exp)
        np.testing.assert_array_equal(gw[..., nbins:], exp[..., :4])
        src = folds[lv]
        wrap_in = torch.cat([src, src[..., :4]], dim=-1).contiguous()
        gi = kernels.ffa_shift_add2(wrap_in, a.table, b.table, a.ncoords, b.ncoords,
                                    in_layout="wrap", out_layout="wrap").cpu().numpy()
        np.testing.assert_array_equal(gi, gw)


def _np_level(fold: np.ndarray, table: np.ndarray) -> np.ndarray:
    """One level in numpy FP32 from a resolve table ([tail row, shift, head row, shift])."""
    nseg, nprev, _, B = fold.shape
    out = np.empty((nseg // 2, len(table), 2, B), dtype=np.float32)
    for q in range(nseg // 2):
        for j, (rt, st, rh, sh) in enumerate(table):
            out[q, j] = np.roll(fold[2 * q, rt], st, axis=-1) + np.roll(fold[2 * q + 1, rh], sh,
                                                                          axis=-1)
    return out


@pytest.mark.parametrize("two", [False, True])
def test_fused_score_signed_zero_profiles(gpu_device, two):
    """Profiles that are zero over whole stretches (+0.0 and -0.0 mixed, as zero-filled gaps
    of a series give): the fused scorers' S/N equals the sequential reference bit for bit,
    the sign of a zero S/N included."""
    _case(f"fused score signed zeros two={two}")
    B, widths = 64, np.array([1, 2, 3, 5, 8, 12])
    rng = np.random.default_rng(3 + two)
    nprev, ncur = 300, 512
    nseg = 4 if two else 2
    e = rng.choice(np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0], dtype=np.float32),
                   size=(nseg, nprev, B))
    e[:, : nprev // 2] = rng.choice(np.array([0.0, -0.0], dtype=np.float32),
                                     size=(nseg, nprev // 2, B))
    fold = np.stack([e, np.ones_like(e)], axis=2).astype(np.float32)
    if two:
        tprev = np.stack([rng.integers(0, nprev, ncur), rng.integers(0, B, ncur),
                          rng.integers(0, nprev, ncur), rng.integers(0, B, ncur)], 1)
        table = np.stack([rng.integers(0, ncur, ncur), rng.integers(0, B, ncur),
                          rng.integers(0, ncur, ncur), rng.integers(0, B, ncur)], 1)
        mid = _np_level(fold, tprev)
        last = _np_level(mid, table)
        got = kernels.ffa_shift_add2_score(_dev(fold), _dev(tprev.astype(np.int32)),
                                           _dev(table.astype(np.int32)), ncur, ncur, widths,
                                           n_inner=ncur)
    else:
        table = np.stack([rng.integers(0, nprev, ncur), rng.integers(0, B, ncur),
                          rng.integers(0, nprev, ncur), rng.integers(0, B, ncur)], 1)
        last = _np_level(fold, table)
        got = kernels.ffa_shift_add_score(_dev(fold), _dev(table.astype(np.int32)), ncur,
                                          widths, n_inner=ncur)
    exp = orc.boxcar_snr(orc.fold_norm(last), widths)
    g = got.cpu().numpy()
    np.testing.assert_array_equal(g.view(np.uint32), exp.view(np.uint32))
    assert np.sum(exp == 0) > 0
