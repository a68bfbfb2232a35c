"""pyl_ffa_level (the entry point a pyloki maintainer binds for unify_fold,
dyn_pffa.py:435-489) called through the C ABI with ctypes, against the oracle's unify_fold:
every level of small FFA searches, P = 1, 2, 3, bit-exact, and the resolve table it
leaves in the caller's workspace against the oracle's rows and shifts."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from pyloki_amd import _lib
from pyloki_amd.config import ParamLimits, PulsarSearchConfig
from pyloki_amd.simulation.pulse import PulseSignalConfig

pytestmark = pytest.mark.gpu


def _case(msg: str) -> None:
    print(f"[case] {msg}", flush=True)


def _cfg(P: int, nsamps: int, nbins: int) -> PulsarSearchConfig:
    if P == 1:
        lim = ParamLimits.from_taylor((195.0, 205.0)).limits
    elif P == 2:
        lim = ParamLimits.from_taylor((195.0, 205.0), accel=(-50.0, 50.0)).limits
    else:
        lim = ParamLimits.from_taylor((330.0, 336.7), accel=(-300.0, 300.0),
                                      jerk=(-200.0, 200.0)).limits
    return PulsarSearchConfig(nsamps=nsamps, tsamp=64e-6, nbins=nbins, eta=1.0, param_limits=lim,
                              bseg_brute=nsamps // 64, use_fourier=False)


def ffa_level_abi(lib, fold_in: torch.Tensor, cfg, level: int, counts_prev, pa,
                  noalloc: bool = False):
    """One level through pyl_ffa_level (or pyl_ffa_level_noalloc): returns (fold, table)."""
    P = len(pa)
    counts_cur = np.array([len(a) for a in pa], dtype=np.int64)
    ncur = int(np.prod(counts_cur))
    nseg, nbins = fold_in.shape[0], fold_in.shape[-1]
    assert fold_in.shape[1] == int(np.prod(counts_prev))
    params = torch.from_numpy(np.concatenate(pa).astype(np.float64)).cuda()
    out = torch.empty((nseg // 2, ncur, 2, nbins), dtype=torch.float32, device="cuda")
    lim = np.ascontiguousarray(cfg.param_limits, dtype=np.float64).reshape(-1)
    cc = np.ascontiguousarray(counts_cur)
    cp = np.ascontiguousarray(counts_prev, dtype=np.int64)
    i64p = ctypes.POINTER(ctypes.c_int64)
    f64p = ctypes.POINTER(ctypes.c_double)
    vp = ctypes.c_void_p
    st = vp(torch.cuda.current_stream().cuda_stream)
    args = (P, vp(params.data_ptr()), cc.ctypes.data_as(i64p), cp.ctypes.data_as(i64p),
            lim.ctypes.data_as(f64p), level, cfg.tseg_brute, nbins, vp(fold_in.data_ptr()), nseg)
    if noalloc:
        rc = lib.pyl_ffa_level_noalloc(*args, vp(out.data_ptr()), st)
        table = None
    else:
        table = torch.empty((ncur, 4), dtype=torch.int32, device="cuda")
        rc = lib.pyl_ffa_level(*args, vp(table.data_ptr()), vp(out.data_ptr()), st)
    if rc != _lib.PYL_OK:
        raise ValueError(lib.pyl_last_error().decode())
    torch.cuda.synchronize()
    return out, table


@pytest.mark.parametrize("P,nbins", [(1, 64), (2, 64), (3, 32), (2, 48)])
def test_pyl_ffa_level_matches_unify_fold(gpu_device, P, nbins):
    _case(f"pyl_ffa_level P={P} nbins={nbins}")
    lib = _lib.load()
    cfg = _cfg(P, 2**16, nbins)
    ts = PulseSignalConfig(period=1 / 200.3, dt=cfg.tsamp, nsamps=cfg.nsamps, snr=20.0,
                           ducy=0.1).generate_analytic(seed=P)
    pa0 = cfg.get_param_arr(cfg.get_dparams(0))
    t_ref = 0.0 if P == 1 else cfg.bseg_brute * cfg.tsamp / 2
    fold = orc.brutefold(ts.ts_e, ts.ts_v, pa0[-1], cfg.bseg_brute, nbins, cfg.tsamp, t_ref)
    fold = fold.reshape(fold.shape[0], -1, 2, nbins)
    dev = torch.from_numpy(fold).cuda()
    counts_prev = cfg.get_param_grid_count(0)
    for level in range(1, cfg.niters_ffa + 1):
        pa = cfg.get_param_arr(cfg.get_dparams(level))
        got, table = ffa_level_abi(lib, dev, cfg, level, counts_prev, pa)
        exp, rows, phases = orc.unify_fold(fold, counts_prev, pa, cfg.param_limits, level,
                                           cfg.tseg_brute, nbins, return_table=True)
        tb = table.cpu().numpy()
        np.testing.assert_array_equal(tb[:, 0], rows[:, 0], err_msg=f"level {level} tail rows")
        np.testing.assert_array_equal(tb[:, 2], rows[:, 1], err_msg=f"level {level} head rows")
        sh = np.array([[orc.shift_int(p, nbins) for p in ph] for ph in phases])
        np.testing.assert_array_equal(tb[:, [1, 3]], sh, err_msg=f"level {level} shifts")
        np.testing.assert_array_equal(got.cpu().numpy(), exp, err_msg=f"level {level}")
        fold, dev = exp, got
        counts_prev = cfg.get_param_grid_count(level)


def test_pyl_ffa_level_rejects_bad_arguments(gpu_device):
    _case("pyl_ffa_level invalid arguments")
    lib = _lib.load()
    cfg = _cfg(2, 2**14, 64)
    pa = cfg.get_param_arr(cfg.get_dparams(1))
    fold = torch.zeros((3, int(np.prod(cfg.get_param_grid_count(0))), 2, 64), device="cuda")
    with pytest.raises(ValueError, match="even"):
        ffa_level_abi(lib, fold, cfg, 1, cfg.get_param_grid_count(0), pa)
