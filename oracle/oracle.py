"""CPU oracle for the pyloki search hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may
import this module, and only as the checker. The product package ``pyloki_amd`` never
imports it.

The numerics live in ``pyl_oracle.c`` (strict IEEE restatement of the reference's numba
kernels, built by ``oracle/Makefile`` into ``oracle/liborc.so``). This module wraps it
with ctypes and adds the level loop of ``DynamicProgramming`` (``ffa.py:305-380``).

How each function is pinned (beyond restating the reference):

- ``phase_idx`` / ``phase_idx_int`` (``psr_utils.py:12-56``): exact rational brute force
  with ``fractions.Fraction`` on random inputs away from rounding boundaries
  (tests/test_oracle.py).
- ``powi`` (``transforms.py:43``, ``psr_utils.py:89``): the binary-powering multiply chain,
  checked against a Python transcription of that chain and, within 1 ulp, against the
  exact power.
- ``brutefold`` (``fold.py:108-171``): a pure-Python loop on tiny inputs, the
  sum-conservation invariant (sum over bins == segment sum, exact for integer-valued
  data), and ``fold_v == per-bin sample count`` for ``ts_v == 1``.
- ``unify_fold`` (``dyn_pffa.py:435-489``): tree replay -- each level-L row equals the
  sum of 2**L level-0 rows rolled by the composed integer shifts (exact for
  integer-valued data), and for P == 1 the FFA fold approximates the direct brute fold.
- ``boxcar_snr_2d`` / ``snr_score_batch`` (``scoring.py:99-165,249-282``): identity with
  the zero-mean unit-energy boxcar matched filter evaluated by brute force over all
  circular shifts (``scoring.py:383-406``), and closed forms for impulse / two-level
  profiles.
- Grid planning is product code (``pyloki_amd.config``), pinned by the notebook
  golden values in tests/golden/.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liborc.so"
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_int = ctypes.c_int


def build() -> Path:
    """Compile the oracle with its Makefile (gcc, strict IEEE)."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(_LIB_PATH))
        L.orc_set_threads.restype = None
        L.orc_set_threads.argtypes = [_int]
        L.orc_get_threads.restype = _int
        L.orc_get_threads.argtypes = []
        L.orc_phase_idx.restype = _dbl
        L.orc_phase_idx.argtypes = [_dbl, _dbl, _i64, _dbl]
        L.orc_phase_idx_int.restype = _i64
        L.orc_phase_idx_int.argtypes = [_dbl, _dbl, _i64, _dbl]
        L.orc_shift_int.restype = _i64
        L.orc_shift_int.argtypes = [_dbl, _i64]
        L.orc_powi.restype = _dbl
        L.orc_powi.argtypes = [_dbl, _int]
        L.orc_nearest_idx.restype = _i64
        L.orc_nearest_idx.argtypes = [_dbl, _i64, _dbl, _dbl]
        L.orc_shift_taylor_params.restype = None
        L.orc_shift_taylor_params.argtypes = [_f64p, _int, _dbl, _f64p]
        L.orc_ffa_resolve.restype = None
        L.orc_ffa_resolve.argtypes = [_int, _f64p, _i64p, _f64p, _int, _int, _dbl, _i64,
                                      _i64p, ctypes.POINTER(_dbl)]
        L.orc_brutefold.restype = None
        L.orc_brutefold.argtypes = [_f32p, _f32p, _i64, _f64p, _i64, _i64, _i64, _dbl, _dbl,
                                    _f32p]
        L.orc_shift_add.restype = None
        L.orc_shift_add.argtypes = [_f32p, _f32p, _dbl, _dbl, _i64, _f32p]
        L.orc_unify_fold.restype = None
        L.orc_unify_fold.argtypes = [_f32p, _i64, _int, _i64p, _f64p, _i64p, _f64p, _int,
                                     _dbl, _i64, _f32p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_boxcar_snr_2d.restype = None
        L.orc_boxcar_snr_2d.argtypes = [_f32p, _i64, _i64, _i64p, _i64, _dbl, _f32p]
        L.orc_snr_score_batch.restype = None
        L.orc_snr_score_batch.argtypes = [_f32p, _i64, _i64, _i64p, _i64, _f32p]
        L.orc_fold_norm.restype = None
        L.orc_fold_norm.argtypes = [_f32p, _i64, _i64, _f32p]
        L.orc_brutefold_f64.restype = None
        L.orc_brutefold_f64.argtypes = [_f32p, _f32p, _i64, _f64p, _i64, _i64, _i64, _dbl,
                                        _dbl, _f64p]
        L.orc_unify_fold_f64.restype = None
        L.orc_unify_fold_f64.argtypes = [_f64p, _i64, _int, _i64p, _f64p, _i64p, _f64p, _int,
                                         _dbl, _i64, _f64p]
        _u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
        L.orc_round_bf16.restype = None
        L.orc_round_bf16.argtypes = [_f32p, _i64, _u16p]
        L.orc_bf16_to_f32_array.restype = None
        L.orc_bf16_to_f32_array.argtypes = [_u16p, _i64, _f32p]
        L.orc_unify_fold_bf16.restype = None
        L.orc_unify_fold_bf16.argtypes = [_u16p, _i64, _int, _i64p, _f64p, _i64p, _f64p, _int,
                                          _dbl, _i64, _u16p]
        L.orc_boxcar_snr_f64.restype = None
        L.orc_boxcar_snr_f64.argtypes = [_f64p, _i64, _i64, _i64p, _i64, _f64p]
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    """Threads of the oracle's parallel loops (the reference's prange loops); results do
    not depend on it."""
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def phase_idx(t: float, f: float, nbins: int, delay: float = 0.0) -> float:
    return lib().orc_phase_idx(t, f, nbins, delay)


def phase_idx_int(t: float, f: float, nbins: int, delay: float = 0.0) -> int:
    return lib().orc_phase_idx_int(t, f, nbins, delay)


def shift_int(x: float, nbins: int) -> int:
    return lib().orc_shift_int(x, nbins)


def powi(x: float, n: int) -> float:
    return lib().orc_powi(x, n)


def shift_taylor_params(vec: np.ndarray, dt: float) -> np.ndarray:
    v = np.ascontiguousarray(vec, dtype=np.float64)
    out = np.empty_like(v)
    lib().orc_shift_taylor_params(v, len(v), dt, out)
    return out


def ffa_resolve(pset, counts_prev, limits, level, latter, tseg_brute, nbins):
    pset = np.ascontiguousarray(pset, dtype=np.float64)
    P = len(pset)
    idx = np.zeros(P, dtype=np.int64)
    ph = _dbl(0.0)
    lib().orc_ffa_resolve(P, pset, np.ascontiguousarray(counts_prev, dtype=np.int64),
                          np.ascontiguousarray(limits, dtype=np.float64).reshape(-1),
                          level, latter, tseg_brute, nbins, idx, ctypes.byref(ph))
    return idx, ph.value


def brutefold(ts_e, ts_v, freqs, seg_len, nbins, tsamp, t_ref=0.0) -> np.ndarray:
    ts_e = np.ascontiguousarray(ts_e, dtype=np.float32)
    ts_v = np.ascontiguousarray(ts_v, dtype=np.float32)
    freqs = np.ascontiguousarray(freqs, dtype=np.float64)
    nseg = -(-len(ts_e) // seg_len)
    out = np.empty((nseg, len(freqs), 2, nbins), dtype=np.float32)
    lib().orc_brutefold(ts_e, ts_v, len(ts_e), freqs, len(freqs), seg_len, nbins, tsamp,
                        t_ref, out)
    return out


def unify_fold(fold_in, counts_prev, param_arr_cur, limits, level, tseg_brute, nbins,
               return_table=False):
    fold_in = np.ascontiguousarray(fold_in, dtype=np.float32)
    nseg_in = fold_in.shape[0]
    P = len(param_arr_cur)
    counts_cur = np.array([len(a) for a in param_arr_cur], dtype=np.int64)
    ncur = int(np.prod(counts_cur))
    params = np.ascontiguousarray(np.concatenate(param_arr_cur), dtype=np.float64)
    out = np.empty((nseg_in // 2, ncur, 2, nbins), dtype=np.float32)
    rows = np.empty((ncur, 2), dtype=np.int64)
    phases = np.empty((ncur, 2), dtype=np.float64)
    lib().orc_unify_fold(fold_in.reshape(-1), nseg_in, P,
                         np.ascontiguousarray(counts_prev, dtype=np.int64), params, counts_cur,
                         np.ascontiguousarray(limits, dtype=np.float64).reshape(-1), level,
                         tseg_brute, nbins, out, rows.ctypes.data, phases.ctypes.data)
    if return_table:
        return out, rows, phases
    return out


def ffa_fold(ts_e, ts_v, cfg, levels=None):
    """Full FFA (``DynamicProgramming.initialize`` + ``execute``, ``ffa.py:305-380``).

    Returns the final fold flattened to ``[nseg, ncoords, 2, nbins]`` and the final
    parameter arrays. ``cfg`` supplies the grids (product planning, golden-pinned).
    """
    nlev = cfg.niters_ffa if levels is None else levels
    pa = cfg.get_param_arr(cfg.get_dparams(0))
    t_ref = 0.0 if cfg.nparams == 1 else cfg.bseg_brute * cfg.tsamp / 2
    fold = brutefold(ts_e, ts_v, pa[-1], cfg.bseg_brute, cfg.nbins, cfg.tsamp, t_ref)
    counts_prev = cfg.get_param_grid_count(0)
    for level in range(1, nlev + 1):
        pa = cfg.get_param_arr(cfg.get_dparams(level))
        fold = unify_fold(fold, counts_prev, pa, cfg.param_limits, level, cfg.tseg_brute,
                          cfg.nbins)
        counts_prev = cfg.get_param_grid_count(level)
    return fold, pa


def fold_norm(folds: np.ndarray) -> np.ndarray:
    f = np.ascontiguousarray(folds, dtype=np.float32)
    nbins = f.shape[-1]
    n = f.size // (2 * nbins)
    out = np.empty((n, nbins), dtype=np.float32)
    lib().orc_fold_norm(f.reshape(-1), n, nbins, out)
    return out.reshape(*f.shape[:-2], nbins)


def boxcar_snr(norm: np.ndarray, widths: np.ndarray, stdnoise: float = 1.0) -> np.ndarray:
    x = np.ascontiguousarray(norm, dtype=np.float32)
    nbins = x.shape[-1]
    n = x.size // nbins
    w = np.ascontiguousarray(widths, dtype=np.int64)
    out = np.empty((n, len(w)), dtype=np.float32)
    lib().orc_boxcar_snr_2d(x.reshape(-1), n, nbins, w, len(w), stdnoise, out)
    return out.reshape(*x.shape[:-1], len(w))


def snr_score_batch(folds: np.ndarray, widths: np.ndarray) -> np.ndarray:
    f = np.ascontiguousarray(folds, dtype=np.float32)
    nbins = f.shape[-1]
    n = f.size // (2 * nbins)
    w = np.ascontiguousarray(widths, dtype=np.int64)
    out = np.empty(n, dtype=np.float32)
    lib().orc_snr_score_batch(f.reshape(-1), n, nbins, w, len(w), out)
    return out


def ffa_search_snr(ts_e, ts_v, cfg) -> np.ndarray:
    """``search.ffa_search`` S/N grid (``search.py:42-47``): score segment 0, per width."""
    fold, pa = ffa_fold(ts_e, ts_v, cfg)
    counts = [len(a) for a in pa]
    norm = fold_norm(fold[0])
    snr = boxcar_snr(norm, cfg.score_widths, 1.0)
    return snr.reshape(*counts, len(cfg.score_widths))


def round_bf16(x: np.ndarray) -> np.ndarray:
    """FP32 -> BF16 bits (uint16), round to nearest even."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, dtype=np.uint16)
    lib().orc_round_bf16(x.reshape(-1), x.size, out.reshape(-1))
    return out


def bf16_to_f32(h: np.ndarray) -> np.ndarray:
    h = np.ascontiguousarray(h, dtype=np.uint16)
    out = np.empty(h.shape, dtype=np.float32)
    lib().orc_bf16_to_f32_array(h.reshape(-1), h.size, out.reshape(-1))
    return out


def ffa_fold_bf16(ts_e, ts_v, cfg) -> tuple[np.ndarray, list]:
    """FFA with BF16 fold storage and FP32 arithmetic (brute fold in FP32, then every
    stored fold element rounded to BF16). Returns BF16 bits [nseg, ncoords, 2, B]."""
    pa = cfg.get_param_arr(cfg.get_dparams(0))
    t_ref = 0.0 if cfg.nparams == 1 else cfg.bseg_brute * cfg.tsamp / 2
    fold = round_bf16(brutefold(ts_e, ts_v, pa[-1], cfg.bseg_brute, cfg.nbins, cfg.tsamp, t_ref))
    counts_prev = np.ascontiguousarray(cfg.get_param_grid_count(0), dtype=np.int64)
    for level in range(1, cfg.niters_ffa + 1):
        pa = cfg.get_param_arr(cfg.get_dparams(level))
        counts_cur = np.array([len(a) for a in pa], dtype=np.int64)
        out = np.empty((fold.shape[0] // 2, int(np.prod(counts_cur)), 2, cfg.nbins),
                       dtype=np.uint16)
        lib().orc_unify_fold_bf16(fold.reshape(-1), fold.shape[0], len(pa), counts_prev,
                                  np.ascontiguousarray(np.concatenate(pa)), counts_cur,
                                  np.ascontiguousarray(cfg.param_limits).reshape(-1), level,
                                  cfg.tseg_brute, cfg.nbins, out.reshape(-1))
        fold = out
        counts_prev = counts_cur
    return fold, pa


def ffa_search_snr_f64(ts_e, ts_v, cfg) -> np.ndarray:
    """FP64 reference of the FFA S/N grid: same resolution (indices, shifts) as the FP32
    path, FP64 data arithmetic throughout. Used to state the error of FP32 and of
    reduced-precision fold storage."""
    L = lib()
    ts_e = np.ascontiguousarray(ts_e, dtype=np.float32)
    ts_v = np.ascontiguousarray(ts_v, dtype=np.float32)
    pa = cfg.get_param_arr(cfg.get_dparams(0))
    t_ref = 0.0 if cfg.nparams == 1 else cfg.bseg_brute * cfg.tsamp / 2
    freqs = np.ascontiguousarray(pa[-1], dtype=np.float64)
    nseg = -(-len(ts_e) // cfg.bseg_brute)
    fold = np.empty((nseg, len(freqs), 2, cfg.nbins), dtype=np.float64)
    L.orc_brutefold_f64(ts_e, ts_v, len(ts_e), freqs, len(freqs), cfg.bseg_brute, cfg.nbins,
                        cfg.tsamp, t_ref, fold.reshape(-1))
    counts_prev = np.ascontiguousarray(cfg.get_param_grid_count(0), dtype=np.int64)
    for level in range(1, cfg.niters_ffa + 1):
        pa = cfg.get_param_arr(cfg.get_dparams(level))
        counts_cur = np.array([len(a) for a in pa], dtype=np.int64)
        out = np.empty((fold.shape[0] // 2, int(np.prod(counts_cur)), 2, cfg.nbins))
        L.orc_unify_fold_f64(fold.reshape(-1), fold.shape[0], len(pa), counts_prev,
                             np.ascontiguousarray(np.concatenate(pa)), counts_cur,
                             np.ascontiguousarray(cfg.param_limits).reshape(-1), level,
                             cfg.tseg_brute, cfg.nbins, out.reshape(-1))
        fold = out
        counts_prev = counts_cur
    w = np.ascontiguousarray(cfg.score_widths, dtype=np.int64)
    n = fold.shape[1]
    snr = np.empty((n, len(w)))
    L.orc_boxcar_snr_f64(np.ascontiguousarray(fold[0]).reshape(-1), n, cfg.nbins, w, len(w),
                         snr.reshape(-1))
    return snr.reshape(*[len(a) for a in pa], len(w))


def cpu_count() -> int:
    return len(os.sched_getaffinity(0))
