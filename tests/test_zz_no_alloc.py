"""pyl_ffa_level_noalloc (the FFA level with no allocation: its resolve table in the
library's fixed device array) against pyl_ffa_level (a caller workspace), through the C
ABI: bit-exact on two calls in a row, the second with a smaller level; a level above the
maximum is rejected before anything is launched."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from pyloki_amd import _lib
from tests.test_gpu_ffa_level_abi import _cfg, ffa_level_abi

pytestmark = pytest.mark.gpu


def _case(msg: str) -> None:
    print(f"[case] {msg}", flush=True)


def _level_input(cfg, seed: int) -> tuple[torch.Tensor, np.ndarray]:
    counts0 = cfg.get_param_grid_count(0)
    rng = np.random.default_rng(seed)
    nseg = cfg.nsamps // cfg.bseg_brute
    f = rng.normal(size=(nseg, int(np.prod(counts0)), 2, cfg.nbins)).astype(np.float32)
    f[:, :, 1] = np.abs(f[:, :, 1]) + 1.0
    return torch.from_numpy(f).cuda(), counts0


def test_noalloc_level_matches_workspace_level(gpu_device):
    _case("pyl_ffa_level_noalloc vs pyl_ffa_level")
    lib = _lib.load()
    cmax = int(lib.pyl_ffa_level_noalloc_max_coords())
    assert cmax == 2**21
    big = _cfg(2, 2**18, 64)
    small = _cfg(3, 2**15, 32)
    for cfg, level, seed in ((big, 3, 1), (small, 2, 2), (big, 3, 3)):
        fold, _ = _level_input(cfg, seed)
        counts_prev = cfg.get_param_grid_count(0)
        if level > 1:  # run the levels before it through the workspace entry point
            for lv in range(1, level):
                pa = cfg.get_param_arr(cfg.get_dparams(lv))
                fold, _ = ffa_level_abi(lib, fold, cfg, lv, counts_prev, pa)
                counts_prev = cfg.get_param_grid_count(lv)
        pa = cfg.get_param_arr(cfg.get_dparams(level))
        ncur = int(np.prod([len(a) for a in pa]))
        assert ncur <= cmax
        exp, _ = ffa_level_abi(lib, fold, cfg, level, counts_prev, pa)
        got, _ = ffa_level_abi(lib, fold, cfg, level, counts_prev, pa, noalloc=True)
        print(f"[result] P={cfg.nparams} level {level}: {ncur} parameter sets, "
              f"{exp.shape[0]} segments: bit-exact", flush=True)
        np.testing.assert_array_equal(got.cpu().numpy(), exp.cpu().numpy())


def test_noalloc_level_rejects_levels_above_the_maximum(gpu_device):
    _case("pyl_ffa_level_noalloc maximum")
    lib = _lib.load()
    cmax = int(lib.pyl_ffa_level_noalloc_max_coords())
    counts = np.array([4, cmax // 2], dtype=np.int64)
    vp = ctypes.c_void_p
    i64p = ctypes.POINTER(ctypes.c_int64)
    rc = lib.pyl_ffa_level_noalloc(2, vp(0), counts.ctypes.data_as(i64p), counts.ctypes.data_as(i64p),
                                   None, 1, 1.0, 64, vp(0), 2, vp(0), vp(0))
    assert rc == _lib.PYL_ERR_INVALID
    assert b"maximum without allocation" in lib.pyl_last_error()
