"""Thin Python wrappers over the C ABI that take and return device tensors.

Each wrapper validates shapes on the host before launching, so that a kernel never runs
on operands it was not written for.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from pyloki_amd import _lib
from pyloki_amd.device import ptr, require_gpu, stream_ptr


class BruteFoldPlan:
    """Device plan of the leaf brute fold for one (freqs, seg_len, nbins, tsamp, t_ref).

    Replaces the per-call phase map of ``brutefold_start`` (``fold.py:150-158``): the map
    is resolved once on the GPU and reused for every series folded with this plan.
    """

    def __init__(self, freqs: np.ndarray, seg_len: int, nbins: int, tsamp: float,
                 t_ref: float) -> None:
        require_gpu()
        lib = _lib.load()
        freqs = np.ascontiguousarray(freqs, dtype=np.float64)
        self.freqs = freqs
        self.nfreq = len(freqs)
        self.seg_len = int(seg_len)
        self.nbins = int(nbins)
        self.tsamp = float(tsamp)
        self.t_ref = float(t_ref)
        handle = ctypes.c_void_p()
        _lib.check(lib.pyl_brutefold_plan_create(
            freqs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), self.nfreq, self.seg_len,
            self.nbins, self.tsamp, self.t_ref, stream_ptr(), ctypes.byref(handle)))
        self._handle = handle

    def fold(self, ts_e: torch.Tensor, ts_v: torch.Tensor, out: torch.Tensor | None = None,
             dtype: torch.dtype = torch.float32) -> torch.Tensor:
        """Fold into FP32 or BF16 storage (FP32 accumulation either way)."""
        if ts_e.shape != ts_v.shape or ts_e.dim() != 1:
            msg = "ts_e, ts_v must be 1-D and of the same length."
            raise ValueError(msg)
        if ts_e.dtype != torch.float32 or ts_v.dtype != torch.float32:
            msg = "ts_e, ts_v must be float32"
            raise ValueError(msg)
        nsamps = ts_e.shape[0]
        nseg = -(-nsamps // self.seg_len)
        shape = (nseg, self.nfreq, 2, self.nbins)
        if out is None:
            out = torch.empty(shape, dtype=dtype, device=ts_e.device)
        elif tuple(out.shape) != shape or out.dtype not in (torch.float32, torch.bfloat16):
            msg = f"out must be float32 or bfloat16 of shape {shape}"
            raise ValueError(msg)
        fn = (_lib.load().pyl_brutefold_execute if out.dtype == torch.float32
              else _lib.load().pyl_brutefold_execute_bf16)
        _lib.check(fn(self._handle, ptr(ts_e), ptr(ts_v), nsamps, ptr(out), stream_ptr()))
        return out

    def fold_batch(self, series: list[tuple[torch.Tensor, torch.Tensor]], out: torch.Tensor,
                   ) -> torch.Tensor:
        """Fold many series (same plan) in one launch per 32 series: series i's rows are
        the next ceil(len_i / seg_len) segments of ``out`` [sum nseg, nfreq, 2, nbins], in
        the given order."""
        n = len(series)
        if n == 0:
            return out
        nseg = 0
        for e, v in series:
            if e.shape != v.shape or e.dim() != 1 or e.dtype != torch.float32 or \
                    v.dtype != torch.float32 or not e.is_contiguous() or not v.is_contiguous():
                msg = "ts_e, ts_v must be contiguous 1-D float32 of the same length."
                raise ValueError(msg)
            nseg += -(-e.shape[0] // self.seg_len)
        shape = (nseg, self.nfreq, 2, self.nbins)
        if tuple(out.shape) != shape or out.dtype not in (torch.float32, torch.bfloat16) or \
                not out.is_contiguous():
            msg = f"out must be contiguous float32 or bfloat16 of shape {shape}"
            raise ValueError(msg)
        pe = (ctypes.c_void_p * n)(*[e.data_ptr() for e, _ in series])
        pv = (ctypes.c_void_p * n)(*[v.data_ptr() for _, v in series])
        lens = _lib.i64_array([e.shape[0] for e, _ in series])
        lib = _lib.load()
        fn = (lib.pyl_brutefold_execute_batch if out.dtype == torch.float32
              else lib.pyl_brutefold_execute_batch_bf16)
        _lib.check(fn(self._handle, ctypes.cast(pe, ctypes.c_void_p),
                      ctypes.cast(pv, ctypes.c_void_p), lens, n, ptr(out), stream_ptr()))
        return out

    @property
    def boundary_count(self) -> int:
        """Phase bins that a +-1 ulp change of their FP64 phase would move."""
        c = ctypes.c_int64(0)
        _lib.check(_lib.load().pyl_brutefold_plan_boundary_count(self._handle, ctypes.byref(c)))
        return int(c.value)

    def __del__(self) -> None:
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                _lib.load().pyl_brutefold_plan_destroy(h)
            except Exception:  # noqa: BLE001 - interpreter shutdown
                pass
            self._handle = None


class ComplexFoldPlan:
    """Device plan of the complex (Fourier-domain) leaf fold (``fold.py:395-481``): the FP32
    basis of the harmonic comb of ``freqs``, built once; ``fold`` is one launch of the
    library's FP32 matrix-core product of the segments and the basis (both components).
    """

    def __init__(self, freqs: np.ndarray, seg_len: int, nbins: int, tsamp: float,
                 t_ref: float) -> None:
        require_gpu()
        lib = _lib.load()
        freqs = np.ascontiguousarray(freqs, dtype=np.float64)
        self.freqs = freqs
        self.nfreq = len(freqs)
        self.seg_len = int(seg_len)
        self.nbins = int(nbins)
        self.nb_f = self.nbins // 2 + 1
        handle = ctypes.c_void_p()
        _lib.check(lib.pyl_cfold_plan_create(
            freqs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), self.nfreq, self.seg_len,
            self.nbins, float(tsamp), float(t_ref), stream_ptr(), ctypes.byref(handle)))
        self._handle = handle

    def fold(self, ts_e: torch.Tensor, ts_v: torch.Tensor,
             out: torch.Tensor | None = None) -> torch.Tensor:
        """complex64 [nseg, nfreq, 2, nbins // 2 + 1]."""
        if ts_e.shape != ts_v.shape or ts_e.dim() != 1:
            msg = "ts_e, ts_v must be 1-D and of the same length."
            raise ValueError(msg)
        if ts_e.dtype != torch.float32 or ts_v.dtype != torch.float32:
            msg = "ts_e, ts_v must be float32"
            raise ValueError(msg)
        nsamps = ts_e.shape[0]
        if nsamps % self.seg_len:
            msg = "the complex fold needs nsamps to be a multiple of segment_len"
            raise ValueError(msg)
        shape = (nsamps // self.seg_len, self.nfreq, 2, self.nb_f)
        if out is None:
            out = torch.empty(shape, dtype=torch.complex64, device=ts_e.device)
        elif tuple(out.shape) != shape or out.dtype != torch.complex64:
            msg = f"out must be complex64 of shape {shape}"
            raise ValueError(msg)
        _lib.check(_lib.load().pyl_cfold_execute(self._handle, ptr(ts_e.contiguous()),
                                                 ptr(ts_v.contiguous()), nsamps, ptr(out),
                                                 stream_ptr()))
        return out

    def __del__(self) -> None:
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                _lib.load().pyl_cfold_plan_destroy(h)
            except Exception:  # noqa: BLE001 - interpreter shutdown
                pass
            self._handle = None


def ffa_shift_add_complex(fold_in: torch.Tensor, table: torch.Tensor, phases: torch.Tensor,
                          ncoords_cur: int, nbins: int,
                          out: torch.Tensor | None = None) -> torch.Tensor:
    """Complex level: [nseg, nprev, 2, nb_f] -> [nseg/2, ncur, 2, nb_f] (complex64)."""
    if fold_in.dim() != 4 or fold_in.shape[2] != 2 or fold_in.dtype != torch.complex64:
        msg = "fold_in must be complex64 [nseg, ncoords, 2, nbins // 2 + 1]"
        raise ValueError(msg)
    nseg, nprev, _, nb_f = fold_in.shape
    if nb_f != nbins // 2 + 1 or nseg % 2:
        msg = "fold_in does not match nbins or has an odd number of segments"
        raise ValueError(msg)
    if tuple(table.shape) != (ncoords_cur, 4) or tuple(phases.shape) != (ncoords_cur, 2):
        msg = "table / phases do not match ncoords_cur"
        raise ValueError(msg)
    shape = (nseg // 2, ncoords_cur, 2, nb_f)
    if out is None:
        out = torch.empty(shape, dtype=torch.complex64, device=fold_in.device)
    _lib.check(_lib.load().pyl_ffa_shift_add_complex(
        ptr(fold_in.contiguous()), nseg, nprev, ptr(table), ptr(phases), ncoords_cur, nbins,
        ptr(out), stream_ptr()))
    return out


def score_boxcar_complex(folds: torch.Tensor, widths: np.ndarray, nbins: int,
                         stdnoise: float = 1.0, mode: int = _lib.PYL_SCORE_PER_WIDTH,
                         out: torch.Tensor | None = None) -> torch.Tensor:
    """Boxcar S/N of complex rows [..., 2, nbins // 2 + 1] (irfft first)."""
    widths = np.ascontiguousarray(widths, dtype=np.int64)
    if folds.dtype != torch.complex64 or folds.shape[-2] != 2 or (
            folds.shape[-1] != nbins // 2 + 1):
        msg = "folds must be complex64 [..., 2, nbins // 2 + 1]"
        raise ValueError(msg)
    lead = tuple(folds.shape[:-2])
    n = int(np.prod(lead)) if lead else 1
    shape = (*lead, len(widths)) if mode == _lib.PYL_SCORE_PER_WIDTH else lead
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=folds.device)
    _lib.check(_lib.load().pyl_score_boxcar_complex(
        ptr(folds.contiguous()), n, nbins, _lib.i64_array(widths), len(widths), float(stdnoise),
        mode, ptr(out), stream_ptr()))
    return out


def ffa_resolve(params_dev: torch.Tensor, counts_cur, counts_prev, param_limits: np.ndarray,
                level: int, tseg_brute: float, nbins: int, table: torch.Tensor | None = None,
                phases: torch.Tensor | None = None) -> torch.Tensor:
    """FP64 resolve of one FFA level -> int32 table [ncoords][4]."""
    P = len(counts_cur)
    ncur = int(np.prod([int(c) for c in counts_cur]))
    if params_dev.dtype != torch.float64 or params_dev.numel() != int(sum(counts_cur)):
        msg = "params_dev must be float64 with sum(counts_cur) entries"
        raise ValueError(msg)
    if table is None:
        table = torch.empty((ncur, 4), dtype=torch.int32, device=params_dev.device)
    if tuple(table.shape) != (ncur, 4) or table.dtype != torch.int32:
        msg = "table must be int32 [ncoords, 4]"
        raise ValueError(msg)
    if phases is not None and (tuple(phases.shape) != (ncur, 2) or phases.dtype != torch.float64):
        msg = "phases must be float64 [ncoords, 2]"
        raise ValueError(msg)
    lim = np.ascontiguousarray(param_limits, dtype=np.float64).reshape(-1)
    _lib.check(_lib.load().pyl_ffa_resolve(
        P, ptr(params_dev), _lib.i64_array(counts_cur), _lib.i64_array(counts_prev),
        _lib.f64_array(lim), level, tseg_brute, nbins, ptr(table), ptr(phases), stream_ptr()))
    return table


FOLD_LAYOUTS = {"dense": _lib.PYL_FOLD_DENSE, "wrap": _lib.PYL_FOLD_WRAP}
WRAP_PAD = 4  # bins copied after each component in the "wrap" row layout


def _layout(name: str) -> int:
    if name not in FOLD_LAYOUTS:
        msg = f"layout must be one of {sorted(FOLD_LAYOUTS)}, got {name!r}"
        raise ValueError(msg)
    return FOLD_LAYOUTS[name]


def _check_level_input(fold_in: torch.Tensor, table: torch.Tensor, ncoords_cur: int,
                       in_layout: str) -> tuple[int, int, int]:
    if fold_in.dim() != 4 or fold_in.shape[2] != 2 or fold_in.dtype not in (
            torch.float32, torch.bfloat16):
        msg = "fold_in must be float32/bfloat16 [nseg, ncoords, 2, nbins]"
        raise ValueError(msg)
    if in_layout != "dense" and fold_in.dtype != torch.float32:
        msg = "the wrap layout is FP32 only"
        raise ValueError(msg)
    nseg, nprev, _, width = fold_in.shape
    nbins = width - (WRAP_PAD if _layout(in_layout) == _lib.PYL_FOLD_WRAP else 0)
    if nseg % 2:
        msg = f"nsegments must be even, got {nseg}"
        raise ValueError(msg)
    if tuple(table.shape) != (ncoords_cur, 4):
        msg = "table does not match ncoords_cur"
        raise ValueError(msg)
    if not fold_in.is_contiguous():
        msg = "fold_in must be contiguous"
        raise ValueError(msg)
    return nseg, nprev, nbins


def ffa_shift_add(fold_in: torch.Tensor, table: torch.Tensor, ncoords_cur: int,
                  out: torch.Tensor | None = None, *, in_layout: str = "dense",
                  out_layout: str = "dense") -> torch.Tensor:
    """Segment-pair shift-add of one level: [nseg, nprev, 2, B] -> [nseg/2, ncur, 2, B].

    FP32 or BF16 storage (FP32 arithmetic); the output has the input's dtype. FP32 folds
    may use the "wrap" row layout ([.., 2, B + 4], each component followed by a copy of
    its first 4 bins) on either side."""
    nseg, nprev, nbins = _check_level_input(fold_in, table, ncoords_cur, in_layout)
    lay_out = _layout(out_layout)
    if out_layout != "dense" and fold_in.dtype != torch.float32:
        msg = "the wrap layout is FP32 only"
        raise ValueError(msg)
    shape = (nseg // 2, ncoords_cur, 2, nbins + (WRAP_PAD if lay_out == _lib.PYL_FOLD_WRAP
                                                  else 0))
    if out is None:
        out = torch.empty(shape, dtype=fold_in.dtype, device=fold_in.device)
    elif tuple(out.shape) != shape or out.dtype != fold_in.dtype or not out.is_contiguous():
        msg = f"out must be contiguous with shape {shape} and dtype {fold_in.dtype}"
        raise ValueError(msg)
    lib = _lib.load()
    if fold_in.dtype == torch.float32:
        _lib.check(lib.pyl_ffa_shift_add_layout(ptr(fold_in), _layout(in_layout), nseg, nprev,
                                                ptr(table), ncoords_cur, nbins, ptr(out),
                                                lay_out, stream_ptr()))
    else:
        _lib.check(lib.pyl_ffa_shift_add_bf16(ptr(fold_in), nseg, nprev, ptr(table),
                                              ncoords_cur, nbins, ptr(out), stream_ptr()))
    return out


def ffa_shift_add_score(fold_in: torch.Tensor, table: torch.Tensor, ncoords_cur: int,
                        widths: np.ndarray, stdnoise: float = 1.0,
                        out: torch.Tensor | None = None, n_inner: int | None = None, *,
                        in_layout: str = "dense") -> torch.Tensor:
    """Last level fused with scoring: [nseg, nprev, 2, B] -> S/N [nseg/2, ncur, n_w].

    ``n_inner`` is the last-axis (frequency) grid count (defaults to ``ncoords_cur``); it
    shapes the output tiling only, not the result. ``in_layout`` as for ffa_shift_add."""
    nseg, nprev, nbins = _check_level_input(fold_in, table, ncoords_cur, in_layout)
    widths = np.ascontiguousarray(widths, dtype=np.int64)
    shape = (nseg // 2, ncoords_cur, len(widths))
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=fold_in.device)
    elif tuple(out.shape) != shape or out.dtype != torch.float32 or not out.is_contiguous():
        msg = f"out must be contiguous float32 with shape {shape}"
        raise ValueError(msg)
    n_inner = ncoords_cur if n_inner is None else int(n_inner)
    lib = _lib.load()
    if fold_in.dtype == torch.float32:
        _lib.check(lib.pyl_ffa_shift_add_score_layout(
            ptr(fold_in), _layout(in_layout), nseg, nprev, ptr(table), ncoords_cur, n_inner,
            nbins, _lib.i64_array(widths), len(widths), float(stdnoise), ptr(out),
            stream_ptr()))
    else:
        _lib.check(lib.pyl_ffa_shift_add_score_bf16(
            ptr(fold_in), nseg, nprev, ptr(table), ncoords_cur, n_inner, nbins,
            _lib.i64_array(widths), len(widths), float(stdnoise), ptr(out), stream_ptr()))
    return out


def ffa_shift_add2(fold_in: torch.Tensor, table_prev: torch.Tensor, table: torch.Tensor,
                   ncoords_prev: int, ncoords_cur: int, out: torch.Tensor | None = None, *,
                   in_layout: str = "dense", out_layout: str = "dense") -> torch.Tensor:
    """Two levels in one pass: [nseg, nprev2, 2, B(+4)] -> [nseg/4, ncur, 2, B(+4)], equal
    bit for bit to ``ffa_shift_add`` with ``table_prev`` then with ``table``; the
    intermediate fold is never stored. FP32, nbins 32..256 (powers of two)."""
    if fold_in.dtype != torch.float32:
        msg = "two fused levels need FP32 folds"
        raise ValueError(msg)
    nseg, nprev2, nbins = _check_level_input(fold_in, table_prev, ncoords_prev, in_layout)
    if nseg % 4:
        msg = f"two fused levels need a multiple of 4 segments, got {nseg}"
        raise ValueError(msg)
    if tuple(table.shape) != (ncoords_cur, 4):
        msg = "table does not match ncoords_cur"
        raise ValueError(msg)
    lay_out = _layout(out_layout)
    shape = (nseg // 4, ncoords_cur, 2, nbins + (WRAP_PAD if lay_out == _lib.PYL_FOLD_WRAP
                                                  else 0))
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=fold_in.device)
    elif tuple(out.shape) != shape or out.dtype != torch.float32 or not out.is_contiguous():
        msg = f"out must be contiguous float32 with shape {shape}"
        raise ValueError(msg)
    _lib.check(_lib.load().pyl_ffa_shift_add2_layout(
        ptr(fold_in), _layout(in_layout), nseg, nprev2, ptr(table_prev), ncoords_prev,
        ptr(table), ncoords_cur, nbins, ptr(out), lay_out, stream_ptr()))
    return out


def ffa_shift_add2_score(fold_in: torch.Tensor, table_prev: torch.Tensor, table: torch.Tensor,
                         ncoords_prev: int, ncoords_cur: int, widths: np.ndarray,
                         stdnoise: float = 1.0, out: torch.Tensor | None = None,
                         n_inner: int | None = None, *, in_layout: str = "dense") -> torch.Tensor:
    """The last two levels fused with scoring: [nseg, nprev2, 2, B] -> S/N
    [nseg/4, ncur, n_w], equal bit for bit to ``ffa_shift_add`` with ``table_prev`` then
    ``ffa_shift_add_score`` with ``table``; the intermediate fold is never stored. FP32,
    nbins 32 or 64, at most 16 widths."""
    if fold_in.dtype != torch.float32:
        msg = "two fused levels need FP32 folds"
        raise ValueError(msg)
    nseg, nprev2, nbins = _check_level_input(fold_in, table_prev, ncoords_prev, in_layout)
    if nseg % 4:
        msg = f"two fused levels need a multiple of 4 segments, got {nseg}"
        raise ValueError(msg)
    if tuple(table.shape) != (ncoords_cur, 4):
        msg = "table does not match ncoords_cur"
        raise ValueError(msg)
    widths = np.ascontiguousarray(widths, dtype=np.int64)
    shape = (nseg // 4, ncoords_cur, len(widths))
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=fold_in.device)
    elif tuple(out.shape) != shape or out.dtype != torch.float32 or not out.is_contiguous():
        msg = f"out must be contiguous float32 with shape {shape}"
        raise ValueError(msg)
    n_inner = ncoords_cur if n_inner is None else int(n_inner)
    lib = _lib.load()
    _lib.check(lib.pyl_ffa_shift_add2_score_layout(
        ptr(fold_in), _layout(in_layout), nseg, nprev2, ptr(table_prev), ncoords_prev,
        ptr(table), ncoords_cur, n_inner, nbins, _lib.i64_array(widths), len(widths),
        float(stdnoise), ptr(out), stream_ptr()))
    return out


def score_boxcar(folds: torch.Tensor, widths: np.ndarray, stdnoise: float = 1.0,
                 layout: int = _lib.PYL_LAYOUT_EV, mode: int = _lib.PYL_SCORE_PER_WIDTH,
                 out: torch.Tensor | None = None) -> torch.Tensor:
    """Boxcar S/N of profiles: EV rows [..., 2, B] or normalised rows [..., B]."""
    widths = np.ascontiguousarray(widths, dtype=np.int64)
    nbins = folds.shape[-1]
    if layout in (_lib.PYL_LAYOUT_EV, _lib.PYL_LAYOUT_EV_BF16):
        if folds.dim() < 2 or folds.shape[-2] != 2:
            msg = "EV layout needs [..., 2, nbins]"
            raise ValueError(msg)
        lead = tuple(folds.shape[:-2])
    else:
        lead = tuple(folds.shape[:-1])
    n = int(np.prod(lead)) if lead else 1
    if folds.dtype == torch.bfloat16 and layout == _lib.PYL_LAYOUT_EV:
        layout = _lib.PYL_LAYOUT_EV_BF16
    elif folds.dtype != torch.float32:
        msg = "folds must be float32 (or bfloat16 (e, v) rows)"
        raise ValueError(msg)
    shape = (*lead, len(widths)) if mode == _lib.PYL_SCORE_PER_WIDTH else lead
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=folds.device)
    _lib.check(_lib.load().pyl_score_boxcar(
        ptr(folds), n, nbins, layout, _lib.i64_array(widths), len(widths), float(stdnoise), mode,
        ptr(out), stream_ptr()))
    return out


def score_boxcar_screen(folds: torch.Tensor, widths: np.ndarray, threshold: float,
                        count: bool = True) -> tuple[torch.Tensor, int | None]:
    """Threshold screen of EV profiles [n, 2, B] (B 32 or 64, stdnoise 1) on the matrix
    cores (``pyl_score_boxcar_screen``): out[i] is the exact maximum S/N of
    ``score_boxcar(..., mode=PYL_SCORE_MAX)`` for every profile that can reach
    ``threshold`` (so ``out >= threshold`` selects exactly the same profiles with the same
    scores), and an upper bound below ``threshold`` for the others. Returns (out, the
    number of profiles scored exactly) (the count read back only with ``count``)."""
    widths = np.ascontiguousarray(widths, dtype=np.int64)
    if folds.dim() != 3 or folds.shape[1] != 2 or folds.dtype != torch.float32:
        msg = "folds must be float32 [n, 2, nbins]"
        raise ValueError(msg)
    n, _, nbins = folds.shape
    out = torch.empty(n, dtype=torch.float32, device=folds.device)
    cnt = torch.zeros(1, dtype=torch.int64, device=folds.device) if count else None
    _lib.check(_lib.load().pyl_score_boxcar_screen(
        ptr(folds.contiguous()), n, nbins, _lib.i64_array(widths), len(widths),
        float(threshold), ptr(out), ptr(cnt), stream_ptr()))
    return out, (int(cnt.item()) if count else None)
