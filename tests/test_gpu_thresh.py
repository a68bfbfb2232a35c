"""GPU parity of the Monte-Carlo threshold design (pyl_thresh_simulate, pyl_rng_normal)
against the oracle (oracle/thresh_oracle.py): the deviates, simulated folds, scores and
survivors bit for bit, and the state tables of evaluate_scheme, determine_scheme and both
modes of DynamicThresholdScheme record for record, for the same seed."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from oracle import thresh_oracle as to
from pyloki_amd import _lib
from pyloki_amd.detection import schemes
from pyloki_amd.detection import thresholding as th
from pyloki_amd.device import ptr, stream_ptr

pytestmark = pytest.mark.gpu


def _case(msg: str) -> None:
    print(f"[case] {msg}", flush=True)


def _states_equal(a: np.ndarray, b: np.ndarray) -> None:
    assert a.dtype == b.dtype and a.shape == b.shape
    for name in a.dtype.names:
        np.testing.assert_array_equal(a[name], b[name], err_msg=name)


@pytest.mark.parametrize("seed,stream,rows,row_len,std", [
    (0, 0, 1000, 64, 1.0), (2**63 + 12345, 2**40 + 7, 300, 100, 2.5), (9, 3, 1, 4, 1.0)])
def test_rng_normal_bit_exact(gpu_device, seed, stream, rows, row_len, std):
    _case(f"rng_normal seed={seed} stream={stream} rows={rows} row_len={row_len}")
    out = torch.empty((rows, row_len), dtype=torch.float32, device=gpu_device)
    _lib.check(_lib.load().pyl_rng_normal(ctypes.c_uint64(seed), ctypes.c_uint64(stream), rows,
                                          row_len, std, ptr(out), stream_ptr()))
    np.testing.assert_array_equal(out.cpu().numpy(),
                                  to.rng_normal(seed, stream, rows, row_len, std))


@pytest.mark.parametrize("nbins,n_prev,ntrials,bias", [
    (64, 300, 1024, 0.0), (64, 1500, 1024, 0.9), (32, 7, 100, 1.3), (16, 64, 64, 0.5)])
def test_simulate_and_score_bit_exact(gpu_device, nbins, n_prev, ntrials, bias):
    _case(f"simulate_folds nbins={nbins} n_prev={n_prev} ntrials={ntrials} bias={bias}")
    rng = np.random.default_rng(nbins + n_prev)
    src = rng.normal(size=(n_prev, nbins)).astype(np.float32)
    prof = th.generate_folded_profile(nbins, 0.1)
    got, var = th.simulate_folds(src, 2.0, prof, (77, 5), bias_snr=bias, var_add=1.0,
                                 ntrials_min=ntrials)
    exp, var_o = to.simulate_folds(src, 2.0, prof * bias, 77, 5, 1.0, ntrials)
    assert var == var_o
    np.testing.assert_array_equal(got, exp)
    for ducy_max, wtsp in ((0.2, 1.0), (0.5, 1.2)):
        sc = th.fold_max_scores(got, var, ducy_max, wtsp)
        np.testing.assert_array_equal(sc, to.fold_max_scores(exp, var, ducy_max, wtsp))
        thr = float(np.median(sc))
        np.testing.assert_array_equal(th.prune_folds(got, var, thr, ducy_max, wtsp),
                                      to.prune_folds(exp, var, thr, ducy_max, wtsp))
        # a threshold equal to a score drops that row (strict >)
        kept = th.prune_folds(got, var, float(sc[0]), ducy_max, wtsp)
        assert len(kept) == int((sc > sc[0]).sum())
    t, rows = th.prune_folds_survival(got, var, 0.3, 0.2, 1.0)
    sc = to.fold_max_scores(exp, var, 0.2, 1.0)
    order = np.argsort(sc, kind="stable")[::-1][: int(0.3 * len(sc))]
    assert t == float(sc[order[-1]])
    np.testing.assert_array_equal(rows, exp[order])


def test_count_mode_matches_compaction(gpu_device):
    """PYL_SIM_COUNT survivor counts for many thresholds equal the number of rows
    PYL_SIM_COMPACT keeps at each, for several tasks in one launch."""
    _case("count vs compact, 6 tasks x 40 thresholds")
    nbins, ntrials = 64, 700
    prof = th.generate_folded_profile(nbins, 0.1)
    sim = th._Sim(nbins, prof, 0.8, 0.3, 1.3, 4242)
    src = torch.from_numpy(np.random.default_rng(1).normal(size=(3 * 500, nbins))
                           .astype(np.float32)).cuda()
    thr = np.linspace(0.5, 6, 40)
    thr_dev = torch.from_numpy(thr).cuda()
    t = th._tasks(6)
    for c in range(6):
        n_prev = 500 - 77 * c
        t[c] = ((c % 3) * 500, 0, c * 1400, c * 40, 1000 + c, 3.0, 0.0, n_prev,
                th._n_sim(n_prev, ntrials), c % 2, 0, 40, 0)
    counts = torch.zeros(240, dtype=torch.int32, device=gpu_device)
    scores = torch.zeros(6 * 1400, dtype=torch.float32, device=gpu_device)
    sim.run(src, t, _lib.PYL_SIM_COUNT, thr=thr_dev, scores=scores, counts=counts)
    cnt = counts.cpu().numpy().reshape(6, 40)
    sc_all = scores.cpu().numpy()
    for c in range(6):
        sc = sc_all[c * 1400: c * 1400 + t[c]["n_sim"]]
        np.testing.assert_array_equal(cnt[c], (sc[:, None] > thr[None, :]).sum(0))
    for it in (0, 17, 39):
        tt = t.copy()
        tt["dst_row"] = np.arange(6) * 1400
        tt["thr"] = thr[it]
        tt["count_off"] = np.arange(6)
        dst = torch.zeros((6 * 1400, nbins), dtype=torch.float32, device=gpu_device)
        kc = torch.zeros(6, dtype=torch.int32, device=gpu_device)
        sim.run(src, tt, _lib.PYL_SIM_COMPACT, dst=dst, counts=kc)
        np.testing.assert_array_equal(kc.cpu().numpy(), cnt[:, it])


@pytest.mark.parametrize("nbins,ntrials,nstages", [(64, 512, 12), (32, 200, 6)])
def test_evaluate_scheme_matches_oracle(gpu_device, nbins, ntrials, nstages):
    _case(f"evaluate_scheme nbins={nbins} ntrials={ntrials} nstages={nstages}")
    rng = np.random.default_rng(nstages)
    bp = rng.choice([1.0, 2.0, 3.0, 1.5], nstages)
    thr = np.sqrt(np.linspace(1, 30, nstages))
    got = th.evaluate_scheme(thr, bp, 0.1, nbins, ntrials, 9.0, 0.3, 1.2, seed=2024)
    exp = to.evaluate_scheme(thr, bp, 0.1, nbins, ntrials, 9.0, 0.3, 1.2, 2024)
    assert len(got.entries) == len(exp)
    for e, r in zip(got.entries, exp):
        assert e == schemes.StateInfo.from_record(r)
    # thresholds above every score: no trial survives, the run raises as the reference's
    with pytest.raises(ZeroDivisionError):
        th.evaluate_scheme(np.full(nstages, 1e6), bp, 0.1, nbins, 64, 9.0, seed=1)


def test_determine_scheme_matches_oracle(gpu_device):
    _case("determine_scheme nbins=64 ntrials=400 nstages=8")
    bp = np.array([3.0, 2.0, 1.0, 3.0, 1.0, 2.0, 1.0, 1.0])
    probs = 1 / bp
    probs[:2] = 1
    got = th.determine_scheme(probs, bp, 0.1, 64, 400, 8.0, 0.2, 1.0, seed=31)
    exp = to.determine_scheme(probs, bp, 0.1, 64, 400, 8.0, 0.2, 1.0, 31)
    assert [schemes.StateInfo.from_record(r) for r in exp] == got.entries


@pytest.mark.parametrize("mode", ["improved", "legacy"])
def test_dynamic_scheme_matches_oracle(gpu_device, mode):
    nstages, nthr, nprobs, ntrials = 6, 20, 5, 128
    _case(f"DynamicThresholdScheme {mode} nstages={nstages} nthr={nthr} nprobs={nprobs}")
    bp = np.array([4.0, 2.0, 3.0, 1.0, 2.0, 1.5])
    dyn = th.DynamicThresholdScheme(bp, 0.1, 64, ntrials=ntrials, nprobs=nprobs, prob_min=0.05,
                                    snr_final=7.0, nthresholds=nthr, ducy_max=0.3, wtsp=1.2,
                                    beam_width=1.5, mode=mode, seed=555)
    dyn.run(thres_neigh=5)
    exp = to.dynamic_scheme(bp, 0.1, 64, ntrials, nprobs, 0.05, 7.0, nthr, 0.3, 1.2, 1.5, mode,
                            555, 5, dyn.guess_path)
    assert (~exp["is_empty"]).sum() > 10
    _states_equal(dyn.states, exp)


def test_dynamic_scheme_save_and_analyse(gpu_device, tmp_path):
    _case("DynamicThresholdScheme save / analyser")
    bp = np.array([3.0, 2.0, 1.0, 2.0, 1.0])
    dyn = th.DynamicThresholdScheme(bp, 0.1, 32, ntrials=96, nprobs=4, snr_final=6.0,
                                    nthresholds=12, beam_width=1.5, mode="improved", seed=8)
    dyn.run(thres_neigh=5)
    path = dyn.save(str(tmp_path))
    assert path.endswith(".npz") and "dynscheme_nstages_005_nthresh_012" in path
    an = schemes.DynamicThresholdSchemeAnalyser.from_file(path)
    _states_equal(an.states, dyn.states)
    assert an.beam_width == 1.5
    best = an.backtrack_best([0.05])[0]
    assert len(best.entries) in (0, 5)
