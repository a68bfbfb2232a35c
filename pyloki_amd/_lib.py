"""ctypes binding of the HIP library ``libpylokihip.so`` (C ABI in include/pyloki_hip.h).

This is the only route to the search kernels: there is no CPU fallback. Loading fails
loudly when the library has not been built, and compute calls fail when no ROCm device
is visible.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libpylokihip.so"

PYL_OK = 0
PYL_ERR_INVALID = -1
PYL_ERR_HIP = -2
PYL_ERR_ALLOC = -3

PYL_LAYOUT_EV = 0
PYL_LAYOUT_NORM = 1
PYL_LAYOUT_EV_BF16 = 2
PYL_FOLD_DENSE = 0
PYL_FOLD_WRAP = 1
PYL_SCORE_PER_WIDTH = 0
PYL_SCORE_MAX = 1
PYL_SIM_COUNT = 0
PYL_SIM_COMPACT = 1
PYL_SIM_NO_NOISE = 1

_c_i32 = ctypes.c_int32
_c_i64 = ctypes.c_int64
_c_dbl = ctypes.c_double
_vp = ctypes.c_void_p
_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)

# (name, restype, argtypes) for every symbol declared in include/pyloki_hip.h.
SIGNATURES: list[tuple[str, object, list[object]]] = [
    ("pyl_last_error", ctypes.c_char_p, []),
    ("pyl_abi_version", ctypes.c_int, []),
    ("pyl_brutefold_plan_create", ctypes.c_int,
     [_f64p, _c_i64, _c_i64, _c_i32, _c_dbl, _c_dbl, _vp, ctypes.POINTER(_vp)]),
    ("pyl_brutefold_plan_destroy", ctypes.c_int, [_vp]),
    ("pyl_brutefold_execute", ctypes.c_int, [_vp, _vp, _vp, _c_i64, _vp, _vp]),
    ("pyl_brutefold_execute_bf16", ctypes.c_int, [_vp, _vp, _vp, _c_i64, _vp, _vp]),
    ("pyl_brutefold_execute_batch", ctypes.c_int, [_vp, _vp, _vp, _i64p, _c_i32, _vp, _vp]),
    ("pyl_brutefold_execute_batch_bf16", ctypes.c_int,
     [_vp, _vp, _vp, _i64p, _c_i32, _vp, _vp]),
    ("pyl_brutefold_plan_boundary_count", ctypes.c_int, [_vp, _i64p]),
    ("pyl_brutefold_set_segments_per_block", ctypes.c_int, [_c_i32]),
    ("pyl_cfold_plan_create", ctypes.c_int,
     [_f64p, _c_i64, _c_i64, _c_i32, _c_dbl, _c_dbl, _vp, ctypes.POINTER(_vp)]),
    ("pyl_cfold_plan_destroy", ctypes.c_int, [_vp]),
    ("pyl_cfold_execute", ctypes.c_int, [_vp, _vp, _vp, _c_i64, _vp, _vp]),
    ("pyl_ffa_shift_add_complex", ctypes.c_int,
     [_vp, _c_i64, _c_i64, _vp, _vp, _c_i64, _c_i32, _vp, _vp]),
    ("pyl_score_boxcar_complex", ctypes.c_int,
     [_vp, _c_i64, _c_i32, _i64p, _c_i32, _c_dbl, _c_i32, _vp, _vp]),
    ("pyl_ffa_shift_add_bf16", ctypes.c_int,
     [_vp, _c_i64, _c_i64, _vp, _c_i64, _c_i32, _vp, _vp]),
    ("pyl_ffa_shift_add_score_bf16", ctypes.c_int,
     [_vp, _c_i64, _c_i64, _vp, _c_i64, _c_i64, _c_i32, _i64p, _c_i32, _c_dbl, _vp, _vp]),
    ("pyl_ffa_resolve", ctypes.c_int,
     [_c_i32, _vp, _i64p, _i64p, _f64p, _c_i32, _c_dbl, _c_i32, _vp, _vp, _vp]),
    ("pyl_ffa_resolve_boundary_count", ctypes.c_int,
     [_c_i32, _vp, _i64p, _i64p, _f64p, _c_i32, _c_dbl, _c_i32, _vp, _vp]),
    ("pyl_ffa_shift_add", ctypes.c_int, [_vp, _c_i64, _c_i64, _vp, _c_i64, _c_i32, _vp, _vp]),
    ("pyl_ffa_set_tile_rows", ctypes.c_int, [_c_i32]),
    ("pyl_ffa_set_packed_tables", ctypes.c_int, [_c_i32]),
    ("pyl_ffa_get_tile_rows", ctypes.c_int, []),
    ("pyl_ffa_level", ctypes.c_int,
     [_c_i32, _vp, _i64p, _i64p, _f64p, _c_i32, _c_dbl, _c_i32, _vp, _c_i64, _vp, _vp, _vp]),
    ("pyl_ffa_level_noalloc", ctypes.c_int,
     [_c_i32, _vp, _i64p, _i64p, _f64p, _c_i32, _c_dbl, _c_i32, _vp, _c_i64, _vp, _vp]),
    ("pyl_ffa_level_noalloc_max_coords", ctypes.c_int64, []),
    ("pyl_ffa_shift_add_score", ctypes.c_int,
     [_vp, _c_i64, _c_i64, _vp, _c_i64, _c_i64, _c_i32, _i64p, _c_i32, _c_dbl, _vp, _vp]),
    ("pyl_ffa_shift_add_layout", ctypes.c_int,
     [_vp, _c_i32, _c_i64, _c_i64, _vp, _c_i64, _c_i32, _vp, _c_i32, _vp]),
    ("pyl_ffa_shift_add_score_layout", ctypes.c_int,
     [_vp, _c_i32, _c_i64, _c_i64, _vp, _c_i64, _c_i64, _c_i32, _i64p, _c_i32, _c_dbl, _vp,
      _vp]),
    ("pyl_ffa_shift_add2_layout", ctypes.c_int,
     [_vp, _c_i32, _c_i64, _c_i64, _vp, _c_i64, _vp, _c_i64, _c_i32, _vp, _c_i32, _vp]),
    ("pyl_ffa_shift_add2_score_layout", ctypes.c_int,
     [_vp, _c_i32, _c_i64, _c_i64, _vp, _c_i64, _vp, _c_i64, _c_i64, _c_i32, _i64p, _c_i32,
      _c_dbl, _vp, _vp]),
    ("pyl_thresh_simulate", ctypes.c_int,
     [_vp, _vp, _c_i32, _c_i32, _vp, _c_dbl, ctypes.c_uint64, _i64p, _c_i32, _vp, _c_i32, _vp,
      _vp, _vp, _vp]),
    ("pyl_rng_normal", ctypes.c_int,
     [ctypes.c_uint64, ctypes.c_uint64, _c_i64, _c_i32, _c_dbl, _vp, _vp]),
    ("pyl_rows_gather", ctypes.c_int, [_vp, _vp, _c_i64, _c_i32, _vp, _vp]),
    ("pyl_bp_counts", ctypes.c_int,
     [_c_i32, _c_i32, _c_i32, _c_i64, _f64p, _f64p, _f64p, _f64p, _f64p, _c_i32, _c_i32, _c_dbl,
      _f64p, _vp]),
    ("pyl_ts_quantiles", ctypes.c_int, [_vp, _c_i64, _c_i64, _f64p, _c_i32, _vp, _vp, _vp]),
    ("pyl_ts_deredden_median", ctypes.c_int, [_vp, _c_i64, _c_i64, _vp, _vp, _vp]),
    ("pyl_ts_normalise", ctypes.c_int, [_vp, _c_i64, _c_i32, _vp, _f64p, _vp]),
    ("pyl_sim_series", ctypes.c_int,
     [_vp, _c_i64, _c_dbl, _c_dbl, _vp, _c_i32, ctypes.c_uint64, ctypes.c_uint64, _c_i32, _vp,
      _vp, _vp, _vp]),
    ("pyl_sim_combine", ctypes.c_int,
     [_vp, _vp, _c_i64, ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp, _vp, _vp]),
    ("pyl_hbm_read", ctypes.c_int, [_vp, _c_i64, _vp, _c_i32, _vp]),
    ("pyl_hbm_copy", ctypes.c_int, [_vp, _vp, _c_i64, _c_i32, _vp]),
    ("pyl_score_boxcar_screen", ctypes.c_int,
     [_vp, _c_i64, _c_i32, _i64p, _c_i32, _c_dbl, _vp, _vp, _vp]),
    ("pyl_score_boxcar", ctypes.c_int,
     [_vp, _c_i64, _c_i32, _c_i32, _i64p, _c_i32, _c_dbl, _c_i32, _vp, _vp]),
    ("pyl_circ_branch", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("pyl_circ_stage1", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    ("pyl_circ_expand", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    ("pyl_circ_children", ctypes.c_int,
     [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("pyl_circ_materialize", ctypes.c_int,
     [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_i64, _vp,
      _vp, _vp, _vp, _vp]),
    ("pyl_circ_ascend", ctypes.c_int,
     [_vp, _vp, _vp, _c_i64, _vp, _c_i64, _i64p, _f64p, _c_i32, _c_dbl, _vp, _vp, _vp]),
    ("pyl_prune_children_complex", ctypes.c_int,
     [_vp, _vp, _vp, _vp, _c_i64, _vp, _vp, _vp, _c_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
      _vp]),
    ("pyl_prune_materialize_complex", ctypes.c_int,
     [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_i64, _vp, _vp,
      _vp, _vp, _vp]),
    ("pyl_prune_ascend_complex", ctypes.c_int,
     [_vp, _vp, _c_i64, _vp, _c_i64, _i64p, _f64p, _c_i32, _c_dbl, _vp, _vp, _vp]),
    ("pyl_circ_children_complex", ctypes.c_int,
     [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("pyl_circ_materialize_complex", ctypes.c_int,
     [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_i64, _vp,
      _vp, _vp, _vp, _vp]),
    ("pyl_circ_ascend_complex", ctypes.c_int,
     [_vp, _vp, _vp, _c_i64, _vp, _c_i64, _i64p, _f64p, _c_i32, _c_dbl, _vp, _vp, _vp]),
    ("pyl_prune_branch", ctypes.c_int, [_vp, _vp, _vp, _c_i64, _vp, _vp, _vp, _vp, _vp]),
    ("pyl_prune_children", ctypes.c_int,
     [_vp, _vp, _vp, _vp, _c_i64, _vp, _vp, _vp, _c_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
      _vp]),
    ("pyl_prune_materialize", ctypes.c_int,
     [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_i64, _vp, _vp,
      _vp, _vp, _vp]),
    ("pyl_prune_ascend", ctypes.c_int,
     [_vp, _vp, _c_i64, _vp, _c_i64, _i64p, _f64p, _c_i32, _c_dbl, _vp, _vp, _vp]),
    ("pyl_cheb_branch", ctypes.c_int, [_vp, _vp, _vp, _c_i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("pyl_ep_timer_create", ctypes.c_int, [_c_i32, ctypes.POINTER(_vp)]),
    ("pyl_ep_timer_destroy", ctypes.c_int, [_vp]),
    ("pyl_ep_timer_read", ctypes.c_int,
     [_vp, ctypes.POINTER(_c_i32), ctypes.POINTER(ctypes.c_float), _c_i32,
      ctypes.POINTER(_c_i32)]),
    ("pyl_cnn_forward_bf16", ctypes.c_int, [_vp, _vp, _c_i64, _c_i32, _c_i32, _vp, _vp, _vp]),
    ("pyl_ep_budget_apply", ctypes.c_int,
     [_vp, _c_i32, _c_i32, _vp, _vp, _vp, _vp, _c_i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("pyl_cnn_features", ctypes.c_int,
     [_vp, _vp, _c_i64, _c_i32, _c_i32, _c_i32, _vp, _vp, _vp, _vp]),
    ("pyl_circ_levels", ctypes.c_int,
     [_vp, _vp, _c_i32, _vp, _f64p, _i64p, _c_i32, _vp, _c_i64, _vp, _vp, _c_i64, _c_i64,
      _c_i32, _vp, _vp, _vp]),
    ("pyl_circ_forest_levels", ctypes.c_int,
     [_vp, _vp, _c_i32, _vp, _f64p, _i64p, _c_i32, _vp, _c_i64, _vp, _vp, _i64p, _c_i64,
      _c_i32, _vp, _vp, _vp, _vp]),
    ("pyl_circ_forest_ascend", ctypes.c_int,
     [_vp, _vp, _c_i32, _vp, _i64p, _c_i32, _vp, _c_i64, _c_i64, _i64p, _f64p, _c_i32,
      _c_dbl, _vp, _vp, _vp]),
    ("pyl_ep_levels", ctypes.c_int,
     [_vp, _c_i32, _vp, _f64p, _i64p, _c_i32, _vp, _c_i64, _vp, _vp, _c_i64, _c_i64, _c_i32,
      _vp, _vp]),
    ("pyl_cnn_num_params", ctypes.c_int, []),
    ("pyl_cnn_cache_floats", ctypes.c_int64, [_c_i32]),
    ("pyl_cnn_forward", ctypes.c_int,
     [_vp, _vp, _c_i64, _c_i32, _c_i32, _vp, _vp, _vp, _vp]),
    ("pyl_cnn_backward", ctypes.c_int,
     [_vp, _vp, _vp, _vp, _c_i64, _c_i32, ctypes.c_float, ctypes.c_float, _vp, _c_i64, _vp,
      _vp]),
    ("pyl_cnn_adam", ctypes.c_int,
     [_vp, _vp, _vp, _vp, _c_i64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
      ctypes.c_float, _c_i32, _vp]),
]
# Chebyshev children / materialize take the Taylor arguments (the parents moved to the
# current domain in place of the leaves); the ascend adds the domain's scale.
_BY_NAME = {name: (rt, args) for name, rt, args in SIGNATURES}
for _suffix in ("", "_complex"):
    for _step in ("children", "materialize"):
        _rt, _args = _BY_NAME[f"pyl_prune_{_step}{_suffix}"]
        SIGNATURES.append((f"pyl_cheb_{_step}{_suffix}", _rt, list(_args)))
    _rt, _args = _BY_NAME[f"pyl_prune_ascend{_suffix}"]
    SIGNATURES.append((f"pyl_cheb_ascend{_suffix}", _rt, [*_args[:9], _c_dbl, *_args[9:]]))

_lib: ctypes.CDLL | None = None


class HipLibraryMissingError(RuntimeError):
    """The HIP library is not built; there is no fallback path."""


def load() -> ctypes.CDLL:
    """Load the library and bind every C-ABI symbol (does not touch the GPU)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            msg = (
                f"{LIB_PATH} is missing: build it with `python -m pyloki_amd.csrc.build` "
                "(pyloki_amd has no CPU fallback)"
            )
            raise HipLibraryMissingError(msg)
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, restype, argtypes in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = lib
    return _lib


def check(rc: int) -> None:
    """Raise the reference's exception type for a non-zero status."""
    if rc == PYL_OK:
        return
    msg = load().pyl_last_error().decode(errors="replace")
    if rc == PYL_ERR_INVALID:
        raise ValueError(msg)
    raise RuntimeError(f"pyloki HIP error {rc}: {msg}")


def i64_array(vals) -> ctypes.Array:
    vals = [int(v) for v in vals]
    return (ctypes.c_int64 * max(len(vals), 1))(*vals)


def f64_array(vals) -> ctypes.Array:
    vals = [float(v) for v in vals]
    return (ctypes.c_double * max(len(vals), 1))(*vals)
