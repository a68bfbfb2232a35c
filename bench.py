"""Throughput benchmark of the FFA search path on one MI355X.

Workload (batched FFA acceleration search, Taylor order 2, FP32): 32 seeded series of
N(0,1) noise with an injected Gaussian pulse (P = 5 ms, accel = +20 m/s^2, 8 % duty),
lengths drawn (seeded) from {2^21, 2^22, 2^23} samples at tsamp = 64 us; accel grid
+-50 m/s^2, freq grid 195-205 Hz, nbins = 64, eta = 1, widths 1..12 (ducy_max 0.2, wtsp 1.5).
All 32 series run per step through one brute fold per series, then one launch per FFA level
over every still-active series (segments concatenated), and the per-width boxcar S/N of
every final trial.

One "step" = the whole batch folded + scored (inputs resident in HBM before timing). The
last level of each series is fused with its scoring, so final folds are never stored.
value = sum over series of final-level trials / wall time per step.

Also reported: the roofline of the dominant kernel (the one with the most time per step),
timed live with HIP events around each of its launches on the launch stream, its HBM
traffic and VALU issue fraction from the committed rocprofv3 counter profile, and the CPU
oracle timed on a bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
VALU_PEAK_TOPS = 39.32  # VALU lane-ops/s (1e12): 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
COUNTER_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "ffa_batch_accel32_counters.md")


def committed_counters(tile_rows: int) -> dict | None:
    """Per-kernel counters of one benchmark step at ``tile_rows`` (VALU and LDS
    instructions, HBM bytes from FETCH_SIZE / WRITE_SIZE scaled by the read/copy probe
    calibration), measured with rocprofv3 PMC passes and committed as a JSON block in the
    counter profile. None when the profile does not cover that tile."""
    try:
        text = open(COUNTER_PROFILE).read()
    except OSError:
        return None
    start = text.find("```json")
    if start < 0:
        return None
    end = text.find("```", start + 7)
    try:
        data = json.loads(text[start + 7:end])
    except ValueError:
        return None
    return data.get("tiles", {}).get(str(tile_rows))


def parse() -> argparse.Namespace:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nseries", type=int, default=32)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="time eager FFA steps instead of replays of the captured step")
    ap.add_argument("--tile-rows", type=int, default=None,
                    help="outer rows per 256-row tile of the fused FFA kernels; default: "
                         "chosen by the autotuner (pyloki_amd/autotune.py) on this device")
    ap.add_argument("--workload", choices=["ffa", "ep-taylor-jerk", "ep-circular-cnn",
                                           "screen"],
                    default="ffa",
                    help="ffa: the headline FFA batch (default); ep-taylor-jerk: the EP "
                         "search of a jerk pulsar (trials/s); ep-circular-cnn: the "
                         "circular-orbit EP with a learned survivor budget vs exhaustive")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the HBM read and copy probes (1 GiB, 4 launches each): "
                         "the counter calibration and the attainable bandwidth")
    return ap.parse_args()


def copy_calibration() -> int:
    """HBM probes of known size (pyl_hbm_read, pyl_hbm_copy): 4 launches each over 1 GiB,
    timed with HIP events. Profiled with FETCH_SIZE / WRITE_SIZE they give the factors that
    turn those counters into bytes; timed, the attainable HBM bandwidth on this device."""
    import torch

    from pyloki_amd import _lib
    from pyloki_amd.device import ptr, stream_ptr

    lib = _lib.load()
    n4 = (1 << 30) // 16
    src = torch.ones(4 * n4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    nblk = 256 * 32
    part = torch.empty(nblk, dtype=torch.float32, device="cuda")
    out = {"bytes_per_launch": n4 * 16, "launches": 4}
    for name, fn in (("read", lambda: lib.pyl_hbm_read(ptr(src), n4, ptr(part), nblk,
                                                        stream_ptr())),
                     ("copy", lambda: lib.pyl_hbm_copy(ptr(src), ptr(dst), n4, nblk,
                                                        stream_ptr()))):
        _lib.check(fn())  # warm-up
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(4):
            _lib.check(fn())
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 4
        moved = n4 * 16 * (2 if name == "copy" else 1)
        out[f"{name}_ms"] = ms
        out[f"{name}_gbs"] = moved / (ms * 1e-3) / 1e9
    print(json.dumps(out))
    return 0


def make_config(nsamps: int):
    from pyloki_amd.config import ParamLimits, PulsarSearchConfig

    lim = ParamLimits.from_taylor((195.0, 205.0), accel=(-50.0, 50.0)).limits
    return PulsarSearchConfig(nsamps=nsamps, tsamp=64e-6, nbins=64, eta=1.0, param_limits=lim,
                              ducy_max=0.2, wtsp=1.5, use_fourier=False)


def make_series(nseries: int):
    from pyloki_amd.simulation.pulse import PulseSignalConfig

    rng = np.random.default_rng(20240601)
    logn = rng.choice([21, 22, 23], size=nseries)
    out = []
    for i, ln in enumerate(logn):
        sim = PulseSignalConfig(period=5e-3, dt=64e-6, nsamps=2 ** int(ln), snr=15.0, ducy=0.08,
                                mod_kwargs={"acc": 20.0})
        out.append(sim.generate_analytic(seed=1000 + i))
    return out


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _all_threads() -> int:
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def cpu_baseline(series, gpu_out, reps: int = 5) -> dict:
    """The oracle (the C restatement; OpenMP over the reference's prange loops: segments,
    parameter sets, profiles) on one 2^22 series of the timed batch: one warm-up, then the
    median of ``reps`` runs, at one thread and at every core this process may use. The
    oracle's S/N grid of that series is also compared with the GPU's grid from the timed
    step (the parity column)."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import oracle as orc

    pick = next((i for i, ts in enumerate(series) if ts.nsamps == 2**22), 0)
    ts = series[pick]
    cfg = make_config(ts.nsamps)
    runs = {}
    snr = None
    for th in (1, _all_threads()):
        orc.set_threads(th)
        snr = orc.ffa_search_snr(ts.ts_e, ts.ts_v, cfg)  # warm-up
        dts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            snr = orc.ffa_search_snr(ts.ts_e, ts.ts_v, cfg)
            dts.append(time.perf_counter() - t0)
        runs[th] = float(np.median(dts))
    trials = snr.size // snr.shape[-1]
    got = gpu_out[pick].cpu().numpy()
    exp = snr.reshape(got.shape)
    th_all = _all_threads()
    return {
        "value": trials / runs[th_all],
        "unit": "trials/s",
        "cores": th_all,
        "kind": "port",
        "sample": f"series {pick} of the batch ({ts.nsamps} samples, {trials} trials, same "
                  f"grids): median of {reps} runs after one warm-up",
        "single_thread": {"value": trials / runs[1], "unit": "trials/s", "cores": 1,
                          "median_s": runs[1]},
        "all_cores_median_s": runs[th_all],
        "nproc": os.cpu_count(),
        "cpu_model": _cpu_model(),
        "parity": {"series": pick, "snr_bit_exact": bool(np.array_equal(got, exp)),
                   "mismatches": int(np.sum(got != exp))},
    }


# Circular-orbit EP with the learned survivor budget: N = 2^24 samples at 64 us, 128 FFA
# segments (127 EP levels), f in [249, 251] Hz, p_orb_min = 6 h, m_c_max = 0.7 Msun
# (x up to 1.5 lt-s), nbins 64; an injected 250 Hz pulsar (5 % duty, S/N 15) in a 6 h
# orbit with x = 1.5 lt-s and a seeded random orbital phase per seed.
EP_CNN = {"nsamps": 2**24, "tsamp": 64e-6, "nseg": 128, "nbins": 64, "freq": (249.0, 251.0),
          "p_orb": 6 * 3600.0, "x_orb": 1.5, "m_c_max": 0.7, "m_p_min": 1.4, "snr": 15.0,
          "thresholds": (0.5, 8.0), "max_sugg": 2**21, "budgets": (2**14, 12288),
          "survey_budget": 12288,
          "train_seeds": tuple(range(1, 9)) + tuple(range(25, 49)), "test_seed": 9, "epochs": 16,
          "batch_seeds": tuple(range(9, 17)), "survey_seeds": tuple(range(9, 25)),
          "calib_seeds": tuple(range(101, 109)), "calib_loss": 0.0}


def ep_circular_dyp(seed: int, w: dict = EP_CNN, use_fourier: bool = True):
    from pyloki_amd.config import ParamLimits, PulsarSearchConfig
    from pyloki_amd.ffa import DynamicProgramming
    from pyloki_amd.simulation.pulse import PulseSignalConfig

    n = w["nsamps"]
    lim = ParamLimits.from_circular(w["freq"], p_orb_min=w["p_orb"], m_c_max=w["m_c_max"],
                                    m_p_min=w["m_p_min"]).limits
    cfg = PulsarSearchConfig(nsamps=n, tsamp=w["tsamp"], nbins=w["nbins"], eta=1.0,
                             param_limits=lim, bseg_brute=n // 1024, bseg_ffa=n // w["nseg"],
                             prune_poly_order=5, p_orb_min=w["p_orb"], m_c_max=w["m_c_max"],
                             m_p_min=w["m_p_min"], ducy_max=0.3, wtsp=1.5, branch_max=256,
                             use_fourier=use_fourier)
    rng = np.random.default_rng(seed)
    f0 = 0.5 * (w["freq"][0] + w["freq"][1]) * (1 + 1e-4 * rng.uniform(-1, 1))
    ts = PulseSignalConfig(period=1 / f0, dt=w["tsamp"], nsamps=n, snr=w["snr"], ducy=0.05,
                           mod_type="circular",
                           mod_kwargs={"p_orb": w["p_orb"], "x_orb": w["x_orb"],
                                       "psi": float(rng.uniform(0, 2 * np.pi))}
                           ).generate_analytic(seed=seed)
    dyp = DynamicProgramming(ts, cfg)
    dyp.initialize()
    dyp.execute()
    return dyp


def ep_circular_cnn() -> int:
    import torch

    from pyloki_amd import cnn

    if not torch.cuda.is_available():
        print("no ROCm device visible", file=sys.stderr)
        return 1
    w = EP_CNN
    thr = np.linspace(w["thresholds"][0], w["thresholds"][1], w["nseg"] - 1)
    # Train on the exhaustive runs of the training seeds; compare on the test seed.
    res = cnn.budget_experiment(ep_circular_dyp, w["train_seeds"], w["test_seed"], thr,
                                w["max_sugg"], w["budgets"], epochs=w["epochs"])
    ranker = res["ranker"]
    # The survey: 16 series of the configuration searched as the 16 trees of one forest
    # (every launch of a level serves all of them): exhaustive vs the learned, hybrid and
    # S/N budgets.
    b = w["survey_budget"]
    # A per-level S/N budget schedule from the exhaustive runs of held-out calibration
    # series: each level's need (the budget that keeps every final candidate's ancestor,
    # calib_loss 0), the recipe (margin, floor, window over levels) chosen by leave-one-out
    # on the calibration series (cnn.select_schedule).
    cal = cnn.calibrate_schedules(ep_circular_dyp, w["calib_seeds"], thr, w["max_sugg"],
                                  loss=w["calib_loss"])
    sel = cnn.select_schedule(cal["needs"]["snr"], cal["survivors"], cal["nlevels"])
    sched = sel["schedule"]
    # The hybrid key only where S/N ranking loses calibration candidates at this budget: up
    # to the last level whose S/N need exceeds it on any calibration series; S/N after.
    net_until = max((lv for nd in cal["needs"]["snr"] for lv, v in nd.items() if v > b),
                    default=-1)
    series = [ep_circular_dyp(s) for s in w["survey_seeds"]]
    hooks = {"exhaustive": None,
             f"hybrid_{b}": lambda: cnn.SurvivorBudget(b, ranker, hybrid=True),
             f"hybrid_{b}_until{net_until}": lambda: cnn.SurvivorBudget(
                 b, ranker, hybrid=True, net_until=net_until),
             f"cnn_{b}": lambda: cnn.SurvivorBudget(b, ranker),
             f"snr_{b}": lambda: cnn.SurvivorBudget(b),
             "snr_schedule": lambda: cnn.SurvivorBudget(sched)}
    survey = cnn.forest_throughput(series, thr, w["max_sugg"], hooks)
    r = survey[f"hybrid_{b}"]
    ex = survey["exhaustive"]
    line = {
        "metric": "EP recall of the exhaustive candidates under a learned survivor budget",
        "value": r["recall"], "unit": "fraction", "n_gpus": 1, "steps": len(series),
        "warmup": 1, "ms_per_step": r["ms_per_search"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (network inference bf16)",
        "data": "synthetic",
        "config": {"workload": "ep_circular_cnn", "nsamps": w["nsamps"], "nseg": w["nseg"],
                   "nbins": w["nbins"], "prune_poly_order": 5, "budget": b,
                   "key": "max(rank by logit, rank by S/N)", "max_sugg": w["max_sugg"],
                   "thresholds": list(w["thresholds"]), "train_seeds": list(w["train_seeds"]),
                   "survey_seeds": list(w["survey_seeds"]), "forest_trees": len(series)},
        "trial_ratio": r["trial_ratio"], "speedup": r["speedup"],
        "survey": survey,
        "exhaustive_trials_per_s": ex["trials"] / len(series) / (ex["ms_per_search"] * 1e-3),
        "schedule": {"calib_seeds": list(w["calib_seeds"]), "loss": w["calib_loss"],
                     "recipe": {k: v for k, v in sel.items() if k != "schedule"},
                     "hybrid_net_until": net_until,
                     "sum": int(sched.sum()), "max": int(sched.max()),
                     "levels_above_survey_budget": int((sched > b).sum())},
        "test_seed": {k: v for k, v in res.items() if k not in ("train", "ranker")},
        "train": res["train"],
    }
    print(json.dumps(line), flush=True)
    return 0


# The boxcar threshold screen on the matrix cores against the exact scorer: 2^22 stored
# profiles of 64 bins (e, v), 3 % with a pulse of random strength, the FFA workload's
# widths (ducy_max 0.2, wtsp 1.5), threshold 6.
def screen_bench(args) -> int:
    import torch

    from pyloki_amd import _lib, kernels
    from pyloki_amd.detection.scoring import generate_box_width_trials
    from pyloki_amd.simulation.pulse import generate_folded_profile

    n, B, thr = 2**22, 64, 6.0
    widths = generate_box_width_trials(B, ducy_max=0.2, wtsp=1.5)
    rng = np.random.default_rng(1)
    e = rng.standard_normal((n, B), dtype=np.float32) * 4.0
    pulsed = np.nonzero(rng.random(n) < 0.03)[0]
    prof = generate_folded_profile(B, 0.1).astype(np.float32)
    e[pulsed] += (rng.uniform(0, 10, len(pulsed))[:, None] * 4.0 * prof[None, :]).astype(
        np.float32)
    f = torch.from_numpy(np.stack([e, np.full_like(e, 16.0)], axis=1)).cuda()
    del e

    def timed(fn, reps=10):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        b.synchronize()
        return a.elapsed_time(b) / reps

    ms_exact = timed(lambda: kernels.score_boxcar(f, widths, mode=_lib.PYL_SCORE_MAX))
    ms_screen = timed(lambda: kernels.score_boxcar_screen(f, widths, thr, count=False))
    exact = kernels.score_boxcar(f, widths, mode=_lib.PYL_SCORE_MAX).cpu().numpy()
    out, nres = kernels.score_boxcar_screen(f, widths, thr)
    out = out.cpu().numpy()
    sel = exact >= thr
    line = {"metric": "profiles scored against a threshold per second", "unit": "profiles/s",
            "value": n / (ms_screen * 1e-3), "n_gpus": 1, "higher_is_better": True,
            "vs_baseline": None, "dtype": "bf16 screen, f32 exact rescoring",
            "data": "synthetic",
            "config": {"workload": "boxcar threshold screen", "profiles": n, "nbins": B,
                       "widths": [int(w) for w in widths], "threshold": thr},
            "exact_profiles_per_s": n / (ms_exact * 1e-3), "ms_exact": ms_exact,
            "ms_screen": ms_screen, "speedup_vs_exact": ms_exact / ms_screen,
            "rescored_fraction": nres / n, "survivors": int(sel.sum()),
            "survivors_equal": bool(np.array_equal(out >= thr, sel)
                                    and np.array_equal(out[sel], exact[sel]))}
    print(json.dumps(line), flush=True)
    return 0


# EP search of a jerk pulsar (Taylor order 3): N = 2^23 samples at 64 us, 64 FFA segments
# (63 EP levels), bseg_brute 1024, nbins 64, eta 1, f in [495, 505] Hz, accel +-30 m/s^2,
# jerk +-2 m/s^3; injected P = 2 ms pulsar (a = 5 m/s^2, j = 0.05 m/s^3, so that its
# acceleration stays inside the grid over the whole observation; 5 % duty, S/N 12); thresholds
# bound_scheme(63, 5); max_sugg 2^22, batches of 1024 parents; reference segment 32.
EP_JERK = {"nsamps": 2**23, "tsamp": 64e-6, "nseg": 64, "nbins": 64, "snr": 12.0,
           "snr_bound": 5.0, "max_sugg": 2**22, "ref_seg": 32, "seed": 7}


def ep_jerk_dyp(w: dict = EP_JERK):
    from pyloki_amd.config import ParamLimits, PulsarSearchConfig
    from pyloki_amd.ffa import DynamicProgramming
    from pyloki_amd.simulation.pulse import PulseSignalConfig

    n = w["nsamps"]
    lim = ParamLimits.from_taylor((495.0, 505.0), accel=(-30.0, 30.0), jerk=(-2.0, 2.0)).limits
    cfg = PulsarSearchConfig(nsamps=n, tsamp=w["tsamp"], nbins=w["nbins"], eta=1.0,
                             param_limits=lim, bseg_brute=1024, bseg_ffa=n // w["nseg"],
                             prune_poly_order=3, ducy_max=0.3, wtsp=1.5, branch_max=64,
                             use_fourier=False)
    ts = PulseSignalConfig(period=0.002, dt=w["tsamp"], nsamps=n, snr=w["snr"], ducy=0.05,
                           mod_kwargs={"acc": 5.0, "jerk": 0.05}).generate_analytic(seed=w["seed"])
    dyp = DynamicProgramming(ts, cfg)
    dyp.initialize()
    dyp.execute()
    return dyp


def _ep_run(pr, ref_seg: int) -> None:
    pr.initialize(ref_seg)
    pr.execute_levels()
    pr._do_ascend()


def ep_taylor_jerk(args) -> int:
    import torch

    from pyloki_amd.detection.schemes import bound_scheme
    from pyloki_amd.prune import Pruning

    if not torch.cuda.is_available():
        print("no ROCm device visible", file=sys.stderr)
        return 1
    w = EP_JERK
    dyp = ep_jerk_dyp()
    thr = bound_scheme(dyp.nsegments - 1, w["snr_bound"]).astype(np.float64)

    def new():
        return Pruning(dyp, thr, max_sugg=w["max_sugg"], batch_size=1024)

    for _ in range(args.warmup):
        _ep_run(new(), w["ref_seg"])
    times, trials = [], 0
    for _ in range(args.steps):
        pr = new()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _ep_run(pr, w["ref_seg"])
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        trials = int(sum(st.n_leaves for st in pr.pstats[1:]))
    # The dominant kernel timed per launch on the launch stream (a separate run).
    pr = new()
    pr.timing = True
    _ep_run(pr, w["ref_seg"])
    kt = pr.kernel_times()
    B, P = w["nbins"], 3
    stats = pr.pstats[1:]
    ch_bytes = 0.0
    for st in stats:
        nb = max(st.n_branches, 1)
        ch_bytes += st.n_leaves * 2 * B * 4 + nb * (2 * B * 4 + 16 * (P + 2)) + st.n_leaves * 24
    ch_ms = float(sum(kt.get("children", [])))
    achieved = ch_bytes / (ch_ms * 1e-3) / 1e9 if ch_ms > 0 else 0.0
    ms = float(np.median(times)) * 1e3
    res = {"n_final": pr.world_tree.size,
           "best_snr": float(pr.world_tree.scores.max()) if pr.world_tree.size else None}
    line = {
        "metric": "EP trials (scored leaves) per second", "value": trials / (ms * 1e-3),
        "unit": "trials/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "ep_taylor_jerk", "nsamps": w["nsamps"], "nseg": w["nseg"],
                   "nbins": B, "prune_poly_order": P, "max_sugg": w["max_sugg"],
                   "thresholds": f"bound_scheme(63, {w['snr_bound']})",
                   "levels": len(stats), "trials_per_step": trials,
                   "max_level_children": max(st.n_leaves for st in stats)},
        "roofline": {"bound": "hbm", "kernel": "prune_children_kernel",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "launches": len(kt.get("children", [])), "kernel_ms_per_step": ch_ms},
        "kernel_ms_per_step": {k: float(sum(v)) for k, v in kt.items()},
        "result": res,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = ep_cpu_baseline(dyp, thr, w)
    print(json.dumps(line), flush=True)
    return 0


def ep_cpu_baseline(dyp, thr, w: dict) -> dict:
    """The oracle (NumPy restatement, one thread) on the first levels of the same search,
    as many as fit about 20 s."""
    from oracle import prune_oracle as po

    cfg = dyp.cfg
    pcfg = {"poly_order": 3, "nbins": cfg.nbins, "eta": cfg.eta, "branch_max": cfg.branch_max,
            "counts_init": np.array([len(a) for a in dyp.param_arr]),
            "param_limits": cfg.param_limits, "widths": cfg.score_widths,
            "tiling": cfg.tiling_strategy}
    t0 = time.perf_counter()
    levels, trials = 0, 0
    for nl in range(1, dyp.nsegments):
        res = po.prune_run(dyp.fold, dyp.param_arr, dyp.dparams_actual, dyp.tseg, thr,
                           w["ref_seg"], pcfg, w["max_sugg"], nlevels=nl)
        levels, trials = nl, int(sum(s.n_leaves for s in res["stats"]))
        if time.perf_counter() - t0 > 10.0:
            break
    t0 = time.perf_counter()
    res = po.prune_run(dyp.fold, dyp.param_arr, dyp.dparams_actual, dyp.tseg, thr,
                       w["ref_seg"], pcfg, w["max_sugg"], nlevels=levels)
    dt = time.perf_counter() - t0
    return {"value": trials / dt, "unit": "trials/s", "cores": 1, "kind": "port",
            "sample": f"the oracle on the first {levels} of 63 levels ({trials} trials)"}


def main() -> int:
    args = parse()
    if args.copy_calibration:
        return copy_calibration()
    if args.workload == "ep-taylor-jerk":
        if args.gpus != 1:
            print("bench.py measures one GPU only", file=sys.stderr)
            return 2
        return ep_taylor_jerk(args)
    if args.workload == "ep-circular-cnn":
        if args.gpus != 1:
            print("bench.py measures one GPU only", file=sys.stderr)
            return 2
        return ep_circular_cnn()
    if args.workload == "screen":
        if args.gpus != 1:
            print("bench.py measures one GPU only", file=sys.stderr)
            return 2
        return screen_bench(args)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the FFA batch "
              "would shard by series across GPUs with no communication.", file=sys.stderr)
        return 2
    import torch

    from pyloki_amd import kernels
    from pyloki_amd.ffa import FFAPlan, ffa_batch

    if not torch.cuda.is_available():
        print("no ROCm device visible", file=sys.stderr)
        return 1
    t_gen = time.perf_counter()
    series = make_series(args.nseries)
    lens = [ts.nsamps for ts in series]
    cfg_max = make_config(max(lens))
    plan = FFAPlan(cfg_max)
    dev_series = [(torch.from_numpy(ts.ts_e).cuda(), torch.from_numpy(ts.ts_v).cuda())
                  for ts in series]
    widths = cfg_max.score_widths
    torch.cuda.synchronize()
    print(f"# setup {time.perf_counter() - t_gen:.1f}s, series lengths "
          f"{sorted(set(lens))} counts {[lens.count(n) for n in sorted(set(lens))]}",
          file=sys.stderr, flush=True)

    # Live timing of every kernel of the step: HIP events around each launch, on the stream
    # it is launched on (torch's current stream, which the wrappers pass down).
    nbins = cfg_max.nbins
    nw = len(widths)
    row = 2 * nbins * 4  # bytes of one FP32 fold row
    # name -> (module attribute, rows out, compulsory bytes, algorithmic bytes) per launch.
    # Compulsory bytes: the input fold read once plus the output written once. Algorithmic
    # bytes: every output row's source rows (2 per level) plus its output.
    two = "ffa_shift_add_score_rounds_kernel<64, two levels>"
    names = {"ffa_shift_add": "ffa_shift_add_kernel<64>",
             "ffa_shift_add2": "ffa_shift_add2_kernel<64>",
             "ffa_shift_add_score": "ffa_shift_add_score_rounds_kernel<64>",
             "ffa_shift_add2_score": two}
    src_rows = {"ffa_shift_add": 2, "ffa_shift_add2": 4, "ffa_shift_add_score": 2,
                "ffa_shift_add2_score": 4}
    events: dict[str, list] = {}
    record = {"on": False}
    origs = {k: getattr(kernels, k) for k in names}

    def _timed(attr):
        fn = origs[attr]

        def wrapper(fold_in, *a, **kw):
            if not record["on"]:
                return fn(fold_in, *a, **kw)
            st = torch.cuda.current_stream()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(st)
            res = fn(fold_in, *a, **kw)
            e1.record(st)
            nrows = res.shape[0] * res.shape[1]
            comp = fold_in.numel() * fold_in.element_size() + res.numel() * res.element_size()
            out_b = res.numel() * res.element_size() // nrows
            algo = nrows * (src_rows[attr] * row + out_b)
            events.setdefault(names[attr], []).append((e0, e1, comp, algo))
            return res
        return wrapper

    for attr in names:
        setattr(kernels, attr, _timed(attr))
    bf_name = "bf_fold_lds_kernel<2, float>"
    orig_bf = plan.bf.fold_batch

    def _timed_bf(b, out):
        if not record["on"]:
            return orig_bf(b, out)
        st = torch.cuda.current_stream()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        res = orig_bf(b, out)
        e1.record(st)
        comp = sum(e.numel() * 8 for e, _ in b) + out.numel() * out.element_size()
        events.setdefault(bf_name, []).append((e0, e1, comp, comp))
        return res

    plan.bf.fold_batch = _timed_bf

    # Trials per step: the final grid of every series (known from the plan, no run needed).
    trials = 0
    for n in lens:
        nlev = int(np.log2(n // cfg_max.bseg_brute))
        trials += plan.levels[nlev - 1].ncoords if nlev > 0 else plan.bf.nfreq
    from pyloki_amd import autotune

    if args.tile_rows is not None:
        autotune.set_tile_rows(args.tile_rows)
        tuned = {"rows": args.tile_rows, "ms": None, "cached": False, "source": "--tile-rows"}
    else:  # time every candidate tile and brute-fold shape on this workload
        sig = (f"ffa_batch/B{nbins}/nw{nw}/lens{sorted(set(lens))}x{len(lens)}/"
               f"grid{[int(c) for c in cfg_max.get_param_grid_count(cfg_max.niters_ffa)]}")
        both = autotune.tune_all(sig, lambda: ffa_batch(plan, dev_series, widths))
        t = both["ffa_tile_rows"]
        tuned = {"rows": t["value"], "ms": t["ms"], "cached": t["cached"],
                 "source": "autotune", "bf_segments_per_block": both["bf_segments_per_block"]}
    for _ in range(args.warmup):
        ffa_batch(plan, dev_series, widths)
    torch.cuda.synchronize()
    # The step captured once as a HIP graph (brute fold, every level, the scoring); the
    # timed steps replay it. --no-graph times eager steps instead.
    step = (lambda: ffa_batch(plan, dev_series, widths))
    if not args.no_graph:
        from pyloki_amd.graph import CapturedFFABatch

        cap = CapturedFFABatch(plan, dev_series, widths)
        step = cap.replay
        for _ in range(args.warmup):
            step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    # Separate instrumented pass for the per-kernel times (events perturb the timed loop).
    nrec = max(1, min(args.steps, 3))
    record["on"] = True
    for _ in range(nrec):
        ffa_batch(plan, dev_series, widths)
    torch.cuda.synchronize()
    record["on"] = False
    for attr, fn in origs.items():
        setattr(kernels, attr, fn)
    plan.bf.fold_batch = orig_bf
    kstats = {}
    for name, evs in events.items():
        tot_ms = sum(ev[0].elapsed_time(ev[1]) for ev in evs)
        n = len(evs)
        kstats[name] = {
            "launches_per_step": n / nrec,
            "ms_per_step": tot_ms / nrec,
            "avg_launch_ms": tot_ms / n,
            "compulsory_bytes_per_launch": sum(ev[2] for ev in evs) / n,
            "algorithmic_bytes_per_launch": sum(ev[3] for ev in evs) / n,
        }
        kstats[name]["compulsory_gbs"] = (kstats[name]["compulsory_bytes_per_launch"]
                                          / (kstats[name]["avg_launch_ms"] * 1e-3) / 1e9)
    dominant = max(kstats, key=lambda k: kstats[k]["ms_per_step"])
    tile = int(tuned["rows"]) if tuned.get("rows") else 16
    counters = committed_counters(tile)
    prof = (counters or {}).get("kernels", {})
    for name, st in kstats.items():
        c = prof.get(name)
        if not c:
            continue
        per_launch = 1.0 / max(c["launches_per_step"], 1)
        if c.get("hbm_bytes_per_step") is not None:
            st["hbm_bytes_per_launch"] = c["hbm_bytes_per_step"] * per_launch
            st["hbm_gbs"] = st["hbm_bytes_per_launch"] / (st["avg_launch_ms"] * 1e-3) / 1e9
            st["hbm_over_compulsory"] = st["hbm_bytes_per_launch"] / st[
                "compulsory_bytes_per_launch"]
        if c.get("valu_insts_per_step") is not None:
            ops = c["valu_insts_per_step"] * 64 / (st["ms_per_step"] * 1e-3) / 1e12
            st["valu_tops"] = ops
            st["valu_frac"] = ops / VALU_PEAK_TOPS
    dk = kstats[dominant]
    ms_per_step = wall / args.steps * 1e3
    if "valu_tops" in dk:
        roof = {"kernel": dominant, "bound": "valu", "achieved": dk["valu_tops"],
                "peak": VALU_PEAK_TOPS, "unit": "TOPS (VALU lane-ops)",
                "frac": dk["valu_frac"]}
    else:  # no counter profile at this tile: the HBM view only
        roof = {"kernel": dominant, "bound": "hbm", "achieved": dk["compulsory_gbs"],
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": dk["compulsory_gbs"] / HBM_PEAK_GBS}
    roof.update({
        "traffic": dk.get("hbm_bytes_per_launch"),
        "hbm_gbs": dk.get("hbm_gbs"),
        "hbm_frac": (dk["hbm_gbs"] / HBM_PEAK_GBS) if "hbm_gbs" in dk else None,
        "compulsory_bytes_per_launch": dk["compulsory_bytes_per_launch"],
        "hbm_over_compulsory": dk.get("hbm_over_compulsory"),
        "avg_launch_ms": dk["avg_launch_ms"],
        "launches_per_step": dk["launches_per_step"],
        "counters": (os.path.relpath(COUNTER_PROFILE, os.path.dirname(os.path.abspath(
            __file__))) + f" (tile {tile})") if counters else None,
        "kernels": kstats,
    })
    line = {
        "metric": "param trials folded+scored/sec & %HBM peak; EP recall vs exhaustive at equal S/N",
        "value": trials / (wall / args.steps),
        "unit": "trials/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded N(0,1) noise + injected accelerated pulse)",
        "config": {
            "workload": "FFA accel search (Taylor order 2), batched: 32 series of 2^21..2^23 "
                        "samples, tsamp 64us, accel +-50 m/s^2, freq 195-205 Hz, nbins 64",
            "series": len(series),
            "samples_total": int(sum(lens)),
            "trials_per_step": trials,
            "final_grid_2^23": [int(c) for c in cfg_max.get_param_grid_count(cfg_max.niters_ffa)],
            "parallelism": "1 GPU; series of one length in one launch per level pair",
            "timed": "eager steps" if args.no_graph else "replays of the step captured as a HIP graph",
            "tile_outer_rows": tuned["rows"],
            "tile_source": tuned["source"],
            "autotune_ms": tuned["ms"],
            "bf_segments_per_block": tuned.get("bf_segments_per_block"),
        },
        "roofline": roof,
        "boundary_evals": plan.boundary_counts(),
    }
    if not args.no_cpu_baseline:
        cb = cpu_baseline(series, res)
        line["parity"] = cb.pop("parity")
        line["cpu_baseline"] = cb
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
